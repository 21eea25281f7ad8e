/*
 * Tokenizer of the structure and parameter readers (reader.c:82-608).
 * Semantics kept from the reference because the parsers depend on them:
 *   - reader_getc advances and returns the new current character; reading
 *     past the end makes it EOF and sets reader_flag_eof;
 *   - a token skips whitespace and whole comments first, then runs up to
 *     whitespace or a comment start; a call made after EOF was reached
 *     fails with reader_err_eof, as does running into EOF before a token;
 *   - reader_skipline stops at '\n' or '\r' and steps past it;
 *   - reader_getcomment returns the text between the next comment start and
 *     the following comment stop.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "reader.h"

static int in_set(const char *set, int c) { return c != EOF && c != 0 && strchr(set, c) != NULL; }

int reader_getc(struct reader *r) {
    if (r->first == r->last) {
        ssize_t n = read(r->fd, r->buff, r->size);
        if (n < 0) return reader_err_io;
        r->last = (int)n;
        r->first = 0;
    }
    if (r->first < r->last)
        r->c = (unsigned char)r->buff[r->first++];
    else {
        r->c = EOF;
        r->flags |= reader_flag_eof;
    }
    if (r->c == '\n' || r->c == '\r') {
        r->line += 1;
        r->col = 0;
    } else
        r->col += 1;
    return r->c;
}

int reader_init(struct reader *r, int fd, const char *ws, const char *comm_start, const char *comm_stop,
                int buffsize) {
    if (r == NULL) return reader_err_null;
    memset(r, 0, sizeof(*r));
    r->fd = fd;
    if ((r->buff = (char *)malloc(buffsize)) == NULL) return reader_err_malloc;
    r->size = buffsize;
    r->line = 1;
    r->ws = ws ? ws : " \f\n\r\t\v";
    r->comm_start = comm_start ? comm_start : "";
    r->comm_stop = comm_stop ? comm_stop : "\n\r";
    if (reader_getc(r) == EOF)
        r->flags = reader_flag_eof;
    else
        r->flags |= reader_flag_ready;
    return reader_err_ok;
}

void reader_close(struct reader *r) {
    if (r == NULL) return;
    free(r->buff);
    r->buff = NULL;
}

/* Whitespace and whole comments before the next token. */
static int skip_blank(struct reader *r) {
    for (;;) {
        if (in_set(r->ws, r->c)) {
            if (reader_getc(r) == EOF) return reader_err_eof;
        } else if (in_set(r->comm_start, r->c)) {
            do {
                if (reader_getc(r) == EOF) return reader_err_eof;
            } while (!in_set(r->comm_stop, r->c));
            if (reader_getc(r) == EOF) return reader_err_eof;
        } else
            return reader_err_ok;
    }
}

int reader_gettoken(struct reader *r, char *buff, int buff_size) {
    int k = 0;
    if (r == NULL || buff == NULL) return reader_err_null;
    if (r->flags & reader_flag_eof) return reader_err_eof;
    if (skip_blank(r) < 0) return reader_err_eof;
    while (r->c != EOF && !in_set(r->ws, r->c) && !in_set(r->comm_start, r->c)) {
        if (k >= buff_size - 1) return reader_err_buff;
        buff[k++] = (char)r->c;
        reader_getc(r);
    }
    buff[k] = 0;
    return k;
}

int reader_skiptoken(struct reader *r) {
    int k = 0;
    if (r == NULL) return reader_err_null;
    if (r->flags & reader_flag_eof) return reader_err_eof;
    if (skip_blank(r) < 0) return reader_err_eof;
    while (r->c != EOF && !in_set(r->ws, r->c) && !in_set(r->comm_start, r->c)) {
        k += 1;
        reader_getc(r);
    }
    return k;
}

int reader_getcomment(struct reader *r, char *buff, int buff_size) {
    int k = 0;
    if (r == NULL || buff == NULL) return reader_err_null;
    if (r->flags & reader_flag_eof) return reader_err_eof;
    while (!in_set(r->comm_start, r->c))
        if (reader_getc(r) == EOF) return reader_err_eof;
    if (reader_getc(r) == EOF) {
        buff[0] = 0;
        return 0;
    }
    while (r->c != EOF && !in_set(r->comm_stop, r->c)) {
        if (k >= buff_size - 1) return reader_err_buff;
        buff[k++] = (char)r->c;
        reader_getc(r);
    }
    buff[k] = 0;
    reader_getc(r);
    return k;
}

int reader_skipline(struct reader *r) {
    int k = 0;
    if (r == NULL) return reader_err_null;
    if (r->flags & reader_flag_eof) return reader_err_eof;
    while (r->c != EOF && r->c != '\n' && r->c != '\r') {
        k += 1;
        reader_getc(r);
    }
    reader_getc(r);
    return k;
}
