/* The partition and selection steps of the wave-parallel quicksort
   (sort_desc_wave in engine_gpu.hip), written as plain loops, against the
   serial quicksort of runner.c:110-174 (descending by the low 16 bits,
   middle pivot, selection sort below 16 elements, explicit stack).  Random
   lists of 1..191 words, a quarter with at most 8 distinct keys; any
   difference in the final order, the overflow result, or a guard of the
   wave version tripping on a valid list fails.  Usage: sort_partition_check
   [count]. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#define STK 16
static int serial(unsigned *a, int N) {
    int lo_s[STK], hi_s[STK], q = 0; lo_s[0] = 0; hi_s[0] = N - 1;
    while (q >= 0) {
        int lo = lo_s[q], hi = hi_s[q]; q--;
        if (hi - lo < 15) {
            for (int i = lo; i < hi; i++) { int im = i; unsigned km = a[i] & 0xffff;
                for (int j = i + 1; j <= hi; j++) { unsigned kj = a[j] & 0xffff; if (kj > km) { im = j; km = kj; } }
                if (im != i) { unsigned t = a[im]; a[im] = a[i]; a[i] = t; } }
        } else {
            unsigned p = a[(lo + hi) / 2] & 0xffff; int i = lo, j = hi;
            while (i <= j) { while ((a[i] & 0xffff) > p) i++; while ((a[j] & 0xffff) < p) j--;
                if (i <= j) { if (i < j) { unsigned t = a[i]; a[i] = a[j]; a[j] = t; } i++; j--; } }
            if (q + 2 >= STK) return -1;
            if (j > (lo + hi) / 2) { if (lo < j) { q++; lo_s[q] = lo; hi_s[q] = j; } if (i < hi) { q++; lo_s[q] = i; hi_s[q] = hi; } }
            else { if (i < hi) { q++; lo_s[q] = i; hi_s[q] = hi; } if (lo < j) { q++; lo_s[q] = lo; hi_s[q] = j; } }
        }
    }
    return 0;
}
static int emulated(unsigned *a, int N) {
    int lo_s[STK], hi_s[STK], q = 0; lo_s[0] = 0; hi_s[0] = N - 1;
    int Ltab[256], Rtab[256];
    int pops = 0;
    while (q >= 0) {
        int lo = lo_s[q], hi = hi_s[q]; q--;
        if (++pops > 2 * N + 2 || lo < 0 || hi >= N || lo > hi) { printf("guard trips N=%d pops=%d\n", N, pops); exit(1); }
        if (hi - lo < 15) {
            int len = hi - lo + 1;
            for (int i = 0; i < len - 1; i++) {   /* max of key<<6 | (63-k) over k >= i */
                unsigned best = 0;
                for (int k = i; k < len; k++) { unsigned pk = ((a[lo + k] & 0xffff) << 6) | (63 - k); if (pk > best) best = pk; }
                int im = 63 - (int)(best & 63);
                unsigned t = a[lo + im]; a[lo + im] = a[lo + i]; a[lo + i] = t;
            }
        } else {
            unsigned p = a[(lo + hi) / 2] & 0xffff;
            int nL = 0, nR = 0, TR = 0;
            for (int x = lo; x <= hi; x++) TR += (a[x] & 0xffff) >= p;
            int seenR = 0;
            for (int x = lo; x <= hi; x++) {
                unsigned k = a[x] & 0xffff;
                if (k <= p) Ltab[++nL] = x;                 /* rank from the left, 1-based */
                if (k >= p) { Rtab[TR - seenR] = x; seenR++; } /* rank from the right */
            }
            nR = TR;
            int K = 0;
            for (int k = 1; k <= nL && k <= nR; k++) if (Ltab[k] <= Rtab[k]) K = k; else break;
            /* check the prefix property used on the GPU (count == last index) */
            int cnt = 0; for (int k = 1; k <= nL && k <= nR; k++) cnt += Ltab[k] <= Rtab[k];
            if (cnt != K) { printf("prefix property fails\n"); exit(1); }
            unsigned tmp[512][2];
            for (int k = 1; k <= K; k++) { tmp[k][0] = a[Ltab[k]]; tmp[k][1] = a[Rtab[k]]; }
            for (int k = 1; k <= K; k++) if (Ltab[k] < Rtab[k]) { a[Ltab[k]] = tmp[k][1]; a[Rtab[k]] = tmp[k][0]; }
            int i = Ltab[K] + 1, j = Rtab[K] - 1;
            if (i <= j) {
                int li = K + 1 <= nL ? Ltab[K + 1] : 1 << 30, rj = K + 1 <= nR ? Rtab[K + 1] : -(1 << 30);
                i = li < Rtab[K] ? li : Rtab[K];
                j = rj > Ltab[K] ? rj : Ltab[K];
            }
            if (q + 2 >= STK) return -1;
            if (j > (lo + hi) / 2) { if (lo < j) { q++; lo_s[q] = lo; hi_s[q] = j; } if (i < hi) { q++; lo_s[q] = i; hi_s[q] = hi; } }
            else { if (i < hi) { q++; lo_s[q] = i; hi_s[q] = hi; } if (lo < j) { q++; lo_s[q] = lo; hi_s[q] = j; } }
        }
    }
    return 0;
}
int main(int argc, char **argv) {
    const long count = argc > 1 ? atol(argv[1]) : 1000000;
    unsigned a[256], b[256];
    srand(12345);
    long tests = 0;
    for (long it = 0; it < count; it++) {
        int N = 1 + rand() % 191;
        int range = 1 + (rand() % 4 == 0 ? rand() % 8 : rand() % 40000);
        for (int i = 0; i < N; i++) { a[i] = ((unsigned)i << 16) | (unsigned)(rand() % range); b[i] = a[i]; }
        int r1 = serial(a, N), r2 = emulated(b, N);
        if (r1 != r2 || (r1 == 0 && memcmp(a, b, sizeof(unsigned) * N))) {
            printf("MISMATCH it=%ld N=%d range=%d r=%d/%d\n", it, N, range, r1, r2); return 1; }
        tests++;
    }
    printf("ok %ld arrays\n", tests);
    return 0;
}
