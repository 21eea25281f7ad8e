/*
 * Error stack with the semantics of the reference's errs.c:61-111: each
 * failing call pushes (id, message, line, function, file); errs_dump prints
 * the stack, most recent first.  Fixed capacity; overflow keeps the oldest
 * entries, as the reference does.
 */
#include <stdio.h>
#include <string.h>

#include "mdcore.h"

#define errs_maxstack 32

struct errs_entry {
    int id, line;
    char msg[160], func[64], file[96];
};

static struct errs_entry errs_stack[errs_maxstack];
static int errs_count = 0;

int errs_register(int id, const char *msg, int line, const char *func, const char *file) {
    if (errs_count < errs_maxstack) {
        struct errs_entry *e = &errs_stack[errs_count++];
        e->id = id;
        e->line = line;
        snprintf(e->msg, sizeof(e->msg), "%s", msg ? msg : "");
        snprintf(e->func, sizeof(e->func), "%s", func ? func : "");
        snprintf(e->file, sizeof(e->file), "%s", file ? file : "");
    }
    return id;
}

int errs_dump(FILE *out) {
    int k;
    if (out == NULL) return -1;
    for (k = errs_count - 1; k >= 0; k--)
        fprintf(out, "%s:%s:%i: %s (%i)\n", errs_stack[k].file, errs_stack[k].func, errs_stack[k].line,
                errs_stack[k].msg, errs_stack[k].id);
    return 0;
}

int errs_gettext(char *buf, int size) {
    int k, pos = 0;
    if (buf == NULL || size <= 0) return -1;
    buf[0] = 0;
    for (k = errs_count - 1; k >= 0 && pos < size - 1; k--) {
        int w = snprintf(buf + pos, size - pos, "%s:%s:%i: %s (%i)\n", errs_stack[k].file, errs_stack[k].func,
                         errs_stack[k].line, errs_stack[k].msg, errs_stack[k].id);
        if (w < 0) break;
        pos += w;
    }
    return errs_count;
}

int errs_clear(void) {
    errs_count = 0;
    return 0;
}
