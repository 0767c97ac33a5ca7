/*
 * ps_ini.h -- INI reader with the reference's parsing rules
 * (parse_ini.c:116-222, :686-847):
 *   - '[Name]' starts a section; only alphanumeric characters of the name count;
 *   - a line whose first character is '#' is a comment, a line of blanks is empty;
 *   - 'key = value': blanks around the key and before the value are dropped,
 *     trailing blanks and the newline are cut; an inline '# ...' stays part of
 *     the value (string values keep it, numbers stop at the blank);
 *   - a key line before any section is ignored; a key line without a value
 *     makes the whole file unreadable;
 *   - the first of duplicate keys wins.
 */
#ifndef PS_INI_H
#define PS_INI_H

#include <stdint.h>

typedef struct ps_ini ps_ini;

/* NULL when the file cannot be opened or parsed */
ps_ini *ps_ini_open(const char *path);
void ps_ini_close(ps_ini *ini);

/* 1 on success (value set), 0 when the section/key is missing or unparsable */
int ps_ini_get_int32(const ps_ini *ini, const char *key, const char *section, int32_t *value);
int ps_ini_get_double(const ps_ini *ini, const char *key, const char *section, double *value);
/* *value is a malloc'ed copy */
int ps_ini_get_string(const ps_ini *ini, const char *key, const char *section, char **value);
/* like get_int32, but a missing key (in an existing section) gives -1
 * (parse_ini.c:321-356) */
int ps_ini_get_int32_ifavail(const ps_ini *ini, const char *key, const char *section, int32_t *value);

#endif
