/* End-to-end timing of the public C ABI, as a C consumer calls it: open a
 * .tbc capture, decode frames with chd_decode_frames_async (a C callback that
 * counts and frees each frame) and with a chd_decode_frame loop, host planes
 * out.  The timing pattern follows the reference's examples/nn_benchmark.c:
 * warm-up kept apart, CLOCK_MONOTONIC, several timed runs, median reported.
 *
 *   chd_abi_bench <file.tbc> <kind> <format> <threads> <passes> <runs>
 *
 * kind: 6, 7 or 8; format: an output_format option value; each async run
 * decodes every frame of the capture `passes` times.  Prints one JSON
 * object. */
#define _POSIX_C_SOURCE 200809L
#include <chromadec/chromadec.h>

#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static atomic_long g_ok, g_fail;

static void on_frame(void *user, chd_status_t s, int64_t idx, chd_frame_t *f)
{
    (void)user;
    (void)idx;
    if (s == CHD_OK) atomic_fetch_add(&g_ok, 1);
    else atomic_fetch_add(&g_fail, 1);
    chd_frame_free(f);
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int cmp_d(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static int die(const char *what, chd_status_t s)
{
    fprintf(stderr, "%s: %s: %s\n", what, chd_status_str(s), chd_last_error());
    return 1;
}

int main(int argc, char **argv)
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s file.tbc kind format threads passes runs\n", argv[0]);
        return 2;
    }
    const int kind = atoi(argv[2]), threads = atoi(argv[4]), passes = atoi(argv[5]), runs = atoi(argv[6]);
    if (passes < 1 || runs < 1 || runs > 64) return 2;
    chd_status_t s = chd_init();
    if (s != CHD_OK) return die("chd_init", s);
    chd_video_t *v = NULL;
    if ((s = chd_video_open_composite(argv[1], NULL, NULL, &v)) != CHD_OK) return die("chd_video_open_composite", s);
    chd_video_info_t info;
    if ((s = chd_video_get_info(v, &info)) != CHD_OK) return die("chd_video_get_info", s);
    chd_decoder_t *d = NULL;
    if ((s = chd_decoder_create(v, (chd_decoder_kind_t)kind, &d)) != CHD_OK) return die("chd_decoder_create", s);
    if ((s = chd_decoder_set_option_str(d, "output_format", argv[3])) != CHD_OK) return die("output_format", s);
    if ((s = chd_decoder_set_option_i32(d, "thread_count", threads)) != CHD_OK) return die("thread_count", s);
    if ((s = chd_decoder_commit(d)) != CHD_OK) return die("chd_decoder_commit", s);

    const int64_t nf = info.num_frames;
    const size_t n = (size_t)nf * (size_t)passes;
    int64_t *idx = malloc(n * sizeof(int64_t));
    for (size_t i = 0; i < n; i++) idx[i] = (int64_t)(i % (size_t)nf);

    /* warm-up: every worker's plan, field cache and pinned frames */
    if ((s = chd_decode_frames_async(d, idx, (size_t)nf, on_frame, NULL, NULL)) != CHD_OK)
        return die("chd_decode_frames_async", s);
    double fps[64];
    for (int r = 0; r < runs; r++) {
        atomic_store(&g_ok, 0);
        const double t0 = now_s();
        if ((s = chd_decode_frames_async(d, idx, n, on_frame, NULL, NULL)) != CHD_OK)
            return die("chd_decode_frames_async", s);
        const double dt = now_s() - t0;
        fps[r] = (double)atomic_load(&g_ok) / dt;
    }
    qsort(fps, (size_t)runs, sizeof(double), cmp_d);
    double mean = 0.0;
    for (int r = 0; r < runs; r++) mean += fps[r] / runs;

    /* synchronous loop: chd_decode_frame then chd_frame_free */
    const int nsync = nf < 32 ? (int)nf : 32;
    chd_frame_t *f = NULL;
    if ((s = chd_decode_frame(d, 0, &f)) != CHD_OK) return die("chd_decode_frame", s);
    chd_frame_free(f);
    const double t1 = now_s();
    for (int i = 0; i < nsync; i++) {
        if ((s = chd_decode_frame(d, i, &f)) != CHD_OK) return die("chd_decode_frame", s);
        chd_frame_free(f);
    }
    const double sync_fps = nsync / (now_s() - t1);

    printf("{\"frames_per_run\": %zu, \"threads\": %d, \"runs\": %d, \"async_median\": %.1f, \"async_mean\": %.1f, "
           "\"async_min\": %.1f, \"async_max\": %.1f, \"sync\": %.1f, \"failed\": %ld}\n",
           n, threads, runs, fps[runs / 2], mean, fps[0], fps[runs - 1], sync_fps, atomic_load(&g_fail));
    chd_decoder_free(d);
    chd_video_free(v);
    free(idx);
    chd_shutdown();
    return 0;
}
