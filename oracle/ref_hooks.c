/*
 * ref_hooks.c -- instrumentation linked into oracle/_ref/libmdsm_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  The reference CPU dedisperser
 * (src/prototypes/cpu_dedisp.c) is compiled unmodified from its own source,
 * with two libc names redirected by -D on the compile line:
 *   malloc       -> mdsm_ref_malloc        (records every block main() allocates:
 *                                          the shift-row pointer array, each
 *                                          shift row, the input, the output)
 *   gettimeofday -> mdsm_ref_gettimeofday  (main() calls it once between
 *                                          generate_data() and the dedispersion
 *                                          loop; on that first call the hook may
 *                                          overwrite the constant 0.1 input with
 *                                          caller-supplied time-major samples)
 * main() never frees its buffers, so after it returns the recorded blocks
 * hold the reference's own shift table and dedispersed output.
 */
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static void **g_ptr;
static size_t *g_size;
static int g_n, g_cap;
static const void *g_src;
static size_t g_src_bytes;
static int g_tod_calls;

void *mdsm_ref_malloc(size_t n)
{
    void *p = malloc(n);
    if (g_n == g_cap) {
        int cap = g_cap ? 2 * g_cap : 1024;
        g_ptr = (void **)realloc(g_ptr, (size_t)cap * sizeof(void *));
        g_size = (size_t *)realloc(g_size, (size_t)cap * sizeof(size_t));
        g_cap = cap;
    }
    g_ptr[g_n] = p;
    g_size[g_n] = n;
    g_n++;
    return p;
}

int mdsm_ref_gettimeofday(struct timeval *tv, void *tz)
{
    if (g_tod_calls++ == 0 && g_src != NULL && g_n >= 2) {
        /* allocation order in main(): ..., input, output */
        void *input = g_ptr[g_n - 2];
        size_t cap = g_size[g_n - 2];
        memcpy(input, g_src, g_src_bytes < cap ? g_src_bytes : cap);
    }
    return gettimeofday(tv, (struct timezone *)tz);
}

void mdsm_ref_hooks_reset(void)
{
    for (int i = 0; i < g_n; i++)
        free(g_ptr[i]);
    g_n = 0;
    g_tod_calls = 0;
    g_src = NULL;
    g_src_bytes = 0;
}

void mdsm_ref_hooks_set_input(const void *src, size_t bytes)
{
    g_src = src;
    g_src_bytes = bytes;
}

int mdsm_ref_hooks_count(void) { return g_n; }
void *mdsm_ref_hooks_ptr(int i) { return (i >= 0 && i < g_n) ? g_ptr[i] : NULL; }
size_t mdsm_ref_hooks_size(int i) { return (i >= 0 && i < g_n) ? g_size[i] : 0; }
