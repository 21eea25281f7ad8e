/*
 * A buffered tokenizer over a file descriptor (reader.h/reader.c of the
 * reference): whitespace-separated tokens, comments from any comm_start
 * character up to a comm_stop character.  Internal to the parameter and
 * structure readers (engine_io.c).
 */
#ifndef MD_READER_H
#define MD_READER_H

#define reader_err_ok     0
#define reader_err_null   -1
#define reader_err_malloc -2
#define reader_err_io     -3
#define reader_err_buff   -4
#define reader_err_eof    -5

#define reader_flag_none  0
#define reader_flag_ready 1
#define reader_flag_eof   2

struct reader {
    unsigned int flags;
    int fd;
    int c; /* current character, EOF at the end */
    char *buff;
    int first, last, size;
    int line, col;
    const char *ws, *comm_start, *comm_stop;
};

int reader_init(struct reader *r, int fd, const char *ws, const char *comm_start, const char *comm_stop,
                int buffsize);
void reader_close(struct reader *r);
int reader_getc(struct reader *r);
int reader_gettoken(struct reader *r, char *buff, int buff_size);
int reader_skiptoken(struct reader *r);
int reader_getcomment(struct reader *r, char *buff, int buff_size);
int reader_skipline(struct reader *r);

#endif
