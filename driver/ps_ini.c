/* ps_ini.c -- see ps_ini.h for the rules (those of parse_ini.c). */
#define _GNU_SOURCE
#include "ps_ini.h"

#include <ctype.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    char *key;
    char *value;
} ini_kv;

typedef struct {
    char *name;
    ini_kv *kv;
    int nkv;
} ini_sec;

struct ps_ini {
    ini_sec *sec;
    int nsec;
};

static char *dup_range(const char *s, size_t n)
{
    char *r = (char *)malloc(n + 1);
    memcpy(r, s, n);
    r[n] = '\0';
    return r;
}

void ps_ini_close(ps_ini *ini)
{
    if (!ini) return;
    for (int i = 0; i < ini->nsec; i++) {
        for (int j = 0; j < ini->sec[i].nkv; j++) {
            free(ini->sec[i].kv[j].key);
            free(ini->sec[i].kv[j].value);
        }
        free(ini->sec[i].kv);
        free(ini->sec[i].name);
    }
    free(ini->sec);
    free(ini);
}

/* '[...]': alphanumeric characters up to ']' or the end of the line */
static char *section_name(const char *line)
{
    size_t n = 0, cap = strlen(line) + 1;
    char *r = (char *)malloc(cap);
    for (const char *p = line + 1; *p && *p != ']'; ++p)
        if (isalnum((unsigned char)*p)) r[n++] = *p;
    r[n] = '\0';
    if (n == 0) {
        free(r);
        return NULL;
    }
    return r;
}

/* 'key = value' -> 0 on success */
static int key_value(const char *line, char **key, char **value)
{
    size_t s = 0;
    while (line[s] == ' ' || line[s] == '\t') s++;
    size_t sep = s;
    while (line[sep] && line[sep] != '=') sep++;
    if (sep == s) return -1;
    size_t e = sep;  /* one past the last key character */
    while (e > s && (line[e - 1] == ' ' || line[e - 1] == '\t')) e--;
    if (line[sep] == '\0') return -1; /* no '=': no value */
    size_t v = sep + 1;
    while (line[v] == ' ' || line[v] == '\t') v++;
    if (line[v] == '\0') return -1;
    size_t ve = strlen(line);
    while (ve > v && (line[ve - 1] == ' ' || line[ve - 1] == '\t' || line[ve - 1] == '\n')) ve--;
    if (ve == v) return -1;
    *key = dup_range(line + s, e - s);
    *value = dup_range(line + v, ve - v);
    return 0;
}

ps_ini *ps_ini_open(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    ps_ini *ini = (ps_ini *)calloc(1, sizeof(ps_ini));
    char *line = NULL;
    size_t cap = 0;
    ssize_t len;
    int ok = 1;
    while (ok && (len = getline(&line, &cap, f)) > 0) {
        size_t i = 0;
        while (line[i] == ' ' || line[i] == '\t') i++;
        if (i == (size_t)len - 1) line[0] = '\0'; /* blanks and the newline only */
        switch (line[0]) {
        case '\0':
        case '#':
            break;
        case '[': {
            char *nm = section_name(line);
            if (!nm) {
                fprintf(stderr, "Could not parse '%s'\n", line);
                ok = 0;
                break;
            }
            ini->sec = (ini_sec *)realloc(ini->sec, sizeof(ini_sec) * (size_t)(ini->nsec + 1));
            ini->sec[ini->nsec].name = nm;
            ini->sec[ini->nsec].kv = NULL;
            ini->sec[ini->nsec].nkv = 0;
            ini->nsec++;
            break;
        }
        default: {
            if (ini->nsec == 0) break; /* before any section: ignored */
            char *k = NULL, *v = NULL;
            if (key_value(line, &k, &v) != 0) {
                fprintf(stderr, "Could not parse '%s'\n", line);
                ok = 0;
                break;
            }
            ini_sec *sc = &ini->sec[ini->nsec - 1];
            sc->kv = (ini_kv *)realloc(sc->kv, sizeof(ini_kv) * (size_t)(sc->nkv + 1));
            sc->kv[sc->nkv].key = k;
            sc->kv[sc->nkv].value = v;
            sc->nkv++;
        }
        }
    }
    free(line);
    fclose(f);
    if (!ok) {
        ps_ini_close(ini);
        return NULL;
    }
    return ini;
}

static const ini_sec *find_sec(const ps_ini *ini, const char *name)
{
    for (int i = 0; i < ini->nsec; i++)
        if (strcmp(ini->sec[i].name, name) == 0) return &ini->sec[i];
    return NULL;
}

static const char *find_val(const ini_sec *s, const char *key)
{
    for (int j = 0; j < s->nkv; j++)
        if (strcmp(s->kv[j].key, key) == 0) return s->kv[j].value;
    return NULL;
}

int ps_ini_get_int32(const ps_ini *ini, const char *key, const char *section, int32_t *value)
{
    const ini_sec *s = find_sec(ini, section);
    if (!s) return 0;
    const char *v = find_val(s, key);
    int32_t t;
    if (!v || sscanf(v, "%" SCNi32, &t) != 1) return 0;
    *value = t;
    return 1;
}

int ps_ini_get_int32_ifavail(const ps_ini *ini, const char *key, const char *section, int32_t *value)
{
    const ini_sec *s = find_sec(ini, section);
    if (!s) return 0;
    const char *v = find_val(s, key);
    if (!v) {
        *value = -1;
        return 1;
    }
    int32_t t;
    if (sscanf(v, "%" SCNi32, &t) != 1) return 0;
    *value = t;
    return 1;
}

int ps_ini_get_double(const ps_ini *ini, const char *key, const char *section, double *value)
{
    const ini_sec *s = find_sec(ini, section);
    if (!s) return 0;
    const char *v = find_val(s, key);
    double t;
    if (!v || sscanf(v, "%lf", &t) != 1) return 0;
    *value = t;
    return 1;
}

int ps_ini_get_string(const ps_ini *ini, const char *key, const char *section, char **value)
{
    const ini_sec *s = find_sec(ini, section);
    if (!s) return 0;
    const char *v = find_val(s, key);
    if (!v) return 0;
    *value = strdup(v);
    return 1;
}
