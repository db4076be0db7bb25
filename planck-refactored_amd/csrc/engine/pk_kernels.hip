// pk_kernels.hip — instantiates the ERI kernels of the classes of shard
// PK_SHARD (compiled once per shard, see the Makefile).
#include "eri_classes.h"

#define PK_DEFINE(a, b, c, d, shard)                                                           \
    PK_DEFINE_IF_##shard(a, b, c, d)

#define PK_DEFINE_NOW(a, b, c, d)                                                              \
    void pk_launch_jk_rys_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) {   \
        pk::launch_jk_rys<double, a, b, c, d, 0>(x, g, b_, s, n);                                  \
    }                                                                                          \
    void pk_launch_jk_os_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) {    \
        pk::launch_jk_os<double, a, b, c, d>(x, g, b_, s, n);                                      \
    }                                                                                          \
    void pk_launch_jk_rys32_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) { \
        pk::launch_jk_rys<float, a, b, c, d, 0>(x, g, b_, s, n);                                   \
    }                                                                                          \
    void pk_launch_jk_os32_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) {  \
        pk::launch_jk_os<float, a, b, c, d>(x, g, b_, s, n);                                       \
    }                                                                                          \
    void pk_launch_jk_os_uhf_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) { \
        pk::launch_jk_os<double, a, b, c, d, 2>(x, g, b_, s, n);                                   \
    }                                                                                          \
    void pk_launch_jk_os_seg_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) { \
        pk::launch_jk_os_seg<double, a, b, c, d>(x, g, b_, s, n);                                  \
    }                                                                                          \
    void pk_launch_jk_os32_seg_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) { \
        pk::launch_jk_os_seg<float, a, b, c, d>(x, g, b_, s, n);                                   \
    }                                                                                          \
    void pk_launch_jk_os_uhf_seg_##a##b##c##d(const pk::JKArgs &x, int g, int b_, hipStream_t s, int &n) { \
        pk::launch_jk_os_seg<double, a, b, c, d, 2>(x, g, b_, s, n);                               \
    }

#define PK_SKIP(a, b, c, d)

#if PK_SHARD == 0
#define PK_DEFINE_IF_0 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_0 PK_SKIP
#endif
#if PK_SHARD == 1
#define PK_DEFINE_IF_1 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_1 PK_SKIP
#endif
#if PK_SHARD == 2
#define PK_DEFINE_IF_2 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_2 PK_SKIP
#endif
#if PK_SHARD == 3
#define PK_DEFINE_IF_3 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_3 PK_SKIP
#endif
#if PK_SHARD == 4
#define PK_DEFINE_IF_4 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_4 PK_SKIP
#endif
#if PK_SHARD == 5
#define PK_DEFINE_IF_5 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_5 PK_SKIP
#endif
#if PK_SHARD == 6
#define PK_DEFINE_IF_6 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_6 PK_SKIP
#endif
#if PK_SHARD == 7
#define PK_DEFINE_IF_7 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_7 PK_SKIP
#endif
#if PK_SHARD == 8
#define PK_DEFINE_IF_8 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_8 PK_SKIP
#endif
#if PK_SHARD == 9
#define PK_DEFINE_IF_9 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_9 PK_SKIP
#endif
#if PK_SHARD == 10
#define PK_DEFINE_IF_10 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_10 PK_SKIP
#endif
#if PK_SHARD == 11
#define PK_DEFINE_IF_11 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_11 PK_SKIP
#endif
#if PK_SHARD == 12
#define PK_DEFINE_IF_12 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_12 PK_SKIP
#endif
#if PK_SHARD == 13
#define PK_DEFINE_IF_13 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_13 PK_SKIP
#endif
#if PK_SHARD == 14
#define PK_DEFINE_IF_14 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_14 PK_SKIP
#endif
#if PK_SHARD == 15
#define PK_DEFINE_IF_15 PK_DEFINE_NOW
#else
#define PK_DEFINE_IF_15 PK_SKIP
#endif

PK_FOR_EACH_CLASS(PK_DEFINE)
