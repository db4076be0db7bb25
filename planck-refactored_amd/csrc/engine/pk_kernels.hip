// pk_kernels.hip — instantiates the ERI kernels of the classes of shard
// PK_SHARD (compiled once per shard, see the Makefile).
#include "eri_classes.h"

#define PK_DEFINE(a, b, c, d, shard)                                                           \
    PK_DEFINE_IF_##shard(a, b, c, d)

#define PK_DEFINE_NOW(a, b, c, d)                                                              \
    void pk_launch_jk_rys_##a##b##c##d(const pk::JKArgs &x, int g, hipStream_t s, int &n) {   \
        pk::launch_jk_rys<a, b, c, d, 0>(x, g, s, n);                                          \
    }                                                                                          \
    void pk_launch_eri_rys_##a##b##c##d(const pk::EriArgs &x, int g, hipStream_t s) {          \
        pk::launch_eri_rys<a, b, c, d, 0>(x, g, s);                                            \
    }                                                                                          \
    void pk_launch_jk_os_##a##b##c##d(const pk::JKArgs &x, int g, hipStream_t s, int &n) {    \
        pk::launch_jk_os<a, b, c, d>(x, g, s, n);                                              \
    }                                                                                          \
    void pk_launch_eri_os_##a##b##c##d(const pk::EriArgs &x, int g, hipStream_t s) {           \
        pk::launch_eri_os<a, b, c, d>(x, g, s);                                                \
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

PK_FOR_EACH_CLASS(PK_DEFINE)
