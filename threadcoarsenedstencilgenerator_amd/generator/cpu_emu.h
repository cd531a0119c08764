// CPU emulation of the gfx950 execution model used by generated STREAM
// kernels (test infrastructure for the generator: it runs the exact emitted
// kernel text on the host).
//
// One call of the kernel function executes one wave: every per-lane value is
// a Lane<T> holding the 64 lanes, so the cross-lane DPP moves are emulated
// exactly (wave_shr:1 / wave_shl:1, with bound_ctrl zeroing or an `old`
// operand).  Buffer loads and stores follow the raw-buffer range check
// measured on gfx950 (stride 0): each element is in range iff its full offset
// (per-lane + wave-uniform SGPR offset) plus its size is within the record
// count; out-of-range elements read 0 and are not stored.  A full offset that
// does not fit in 32 bits is counted in tcsg_emu_bad_access (its hardware
// meaning would depend on wrap-around); tests require the count to stay 0.
#pragma once
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <pthread.h>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(x)

namespace tcsg_emu {
constexpr int W = 64;
}

template <class T>
struct Lane {
    T v[tcsg_emu::W];
    Lane() { for (auto &x : v) x = T(); }
    template <class U, class = std::enable_if_t<std::is_arithmetic<U>::value>>
    Lane(U s) { for (auto &x : v) x = (T)s; }
    template <class U>
    explicit Lane(const Lane<U> &o) { for (int i = 0; i < tcsg_emu::W; ++i) v[i] = (T)o.v[i]; }
    T &operator[](int i) { return v[i]; }
    const T &operator[](int i) const { return v[i]; }
};

using u32 = Lane<uint32_t>;
using lbool = Lane<uint8_t>;

#define TCSG_BIN(op)                                                                               \
    template <class A, class B>                                                                    \
    auto operator op(const Lane<A> &a, const Lane<B> &b)->Lane<decltype(A() op B())> {             \
        Lane<decltype(A() op B())> r;                                                              \
        for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = a.v[i] op b.v[i];                           \
        return r;                                                                                  \
    }                                                                                              \
    template <class A, class B, class = std::enable_if_t<std::is_arithmetic<B>::value>>            \
    auto operator op(const Lane<A> &a, B b)->Lane<decltype(A() op B())> {                          \
        Lane<decltype(A() op B())> r;                                                              \
        for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = a.v[i] op b;                                \
        return r;                                                                                  \
    }                                                                                              \
    template <class A, class B, class = std::enable_if_t<std::is_arithmetic<A>::value>>            \
    auto operator op(A a, const Lane<B> &b)->Lane<decltype(A() op B())> {                          \
        Lane<decltype(A() op B())> r;                                                              \
        for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = a op b.v[i];                                \
        return r;                                                                                  \
    }
TCSG_BIN(+)
TCSG_BIN(-)
TCSG_BIN(*)
TCSG_BIN(/)
TCSG_BIN(%)
TCSG_BIN(&)
TCSG_BIN(>>)
TCSG_BIN(<<)
#undef TCSG_BIN

#define TCSG_CMP(op)                                                                               \
    template <class A, class B>                                                                    \
    lbool operator op(const Lane<A> &a, const Lane<B> &b) {                                        \
        lbool r;                                                                                   \
        for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = a.v[i] op b.v[i];                           \
        return r;                                                                                  \
    }                                                                                              \
    template <class A, class B, class = std::enable_if_t<std::is_arithmetic<B>::value>>            \
    lbool operator op(const Lane<A> &a, B b) {                                                     \
        lbool r;                                                                                   \
        for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = a.v[i] op b;                                \
        return r;                                                                                  \
    }
TCSG_CMP(<)
TCSG_CMP(<=)
TCSG_CMP(>)
TCSG_CMP(>=)
TCSG_CMP(==)
TCSG_CMP(&&)
#undef TCSG_CMP

template <class T, class A, class B>
Lane<T> tcsg_sel_impl(const lbool &c, const A &a, const B &b)
{
    Lane<T> la(a), lb(b), r;
    for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = c.v[i] ? la.v[i] : lb.v[i];
    return r;
}
template <class T>
Lane<T> tcsg_sel(const lbool &c, const Lane<T> &a, const Lane<T> &b) { return tcsg_sel_impl<T>(c, a, b); }
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
Lane<T> tcsg_sel(const lbool &c, const Lane<T> &a, S b) { return tcsg_sel_impl<T>(c, a, Lane<T>(b)); }
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
Lane<T> tcsg_sel(const lbool &c, S a, const Lane<T> &b) { return tcsg_sel_impl<T>(c, Lane<T>(a), b); }

inline u32 tcsg_min(const u32 &a, const u32 &b)
{
    u32 r;
    for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = a.v[i] < b.v[i] ? a.v[i] : b.v[i];
    return r;
}

template <class T>
Lane<T> tcsg_fma(const Lane<T> &a, const Lane<T> &b, const Lane<T> &c)
{
    Lane<T> r;
    for (int i = 0; i < tcsg_emu::W; ++i) r.v[i] = std::fma(a.v[i], b.v[i], c.v[i]);
    return r;
}

// DPP wave_shr:1 (lane i <- lane i-1) and wave_shl:1 (lane i <- lane i+1).
template <class T>
Lane<T> tcsg_shr1(const Lane<T> &v)
{
    Lane<T> r(0);
    for (int i = 1; i < tcsg_emu::W; ++i) r.v[i] = v.v[i - 1];
    return r;
}
template <class T>
Lane<T> tcsg_shl1(const Lane<T> &v)
{
    Lane<T> r(0);
    for (int i = 0; i + 1 < tcsg_emu::W; ++i) r.v[i] = v.v[i + 1];
    return r;
}
template <class T>
Lane<T> tcsg_shr1_or(const Lane<T> &old, const Lane<T> &v)
{
    Lane<T> r = tcsg_shr1(v);
    r.v[0] = old.v[0];
    return r;
}
template <class T>
Lane<T> tcsg_shl1_or(const Lane<T> &old, const Lane<T> &v)
{
    Lane<T> r = tcsg_shl1(v);
    r.v[tcsg_emu::W - 1] = old.v[tcsg_emu::W - 1];
    return r;
}

// DPP row_shl:n (lane i <- lane i+n) and row_shr:n (lane i <- lane i-n), each within its
// row of 16 lanes; lanes whose source lies outside the row get 0.
template <int N, class T>
Lane<T> tcsg_emu_rowshl(const Lane<T> &v)
{
    Lane<T> r(0);
    for (int i = 0; i < tcsg_emu::W; ++i)
        if ((i & 15) + N < 16) r.v[i] = v.v[i + N];
    return r;
}
template <int N, class T>
Lane<T> tcsg_emu_rowshr(const Lane<T> &v)
{
    Lane<T> r(0);
    for (int i = 0; i < tcsg_emu::W; ++i)
        if ((i & 15) >= N) r.v[i] = v.v[i - N];
    return r;
}

struct tcsg_emu_rsrc {
    char *base;
    uint32_t records;
};
inline tcsg_emu_rsrc tcsg_make_rsrc(const void *p, uint32_t bytes) { return {(char *)p, bytes}; }

extern "C" {
extern unsigned long tcsg_emu_bad_access;
}
// counted atomically: the waves of an LDS-staging kernel's workgroup run as host threads
inline void tcsg_emu_bad() { __atomic_fetch_add(&tcsg_emu_bad_access, 1ul, __ATOMIC_RELAXED); }

// A value the hardware holds in scalar registers (wave-uniform): lane 0's value; a lane
// that disagrees is counted in tcsg_emu_bad_access.
template <class T>
T tcsg_uni(const Lane<T> &v)
{
    for (int i = 1; i < tcsg_emu::W; ++i)
        if (v.v[i] != v.v[0]) tcsg_emu_bad();
    return v.v[0];
}
template <class T, class = std::enable_if_t<std::is_arithmetic<T>::value>>
T tcsg_uni(T v) { return v; }

template <class T>
Lane<T> tcsg_emu_ld(const tcsg_emu_rsrc &r, const u32 &off, const u32 &soff)
{
    Lane<T> out(0);
    for (int i = 0; i < tcsg_emu::W; ++i) {
        const uint64_t full = (uint64_t)off.v[i] + soff.v[i];
        if (full >> 32) { tcsg_emu_bad(); continue; }
        if (full + sizeof(T) > r.records) continue;  // out of range: reads 0
        std::memcpy(&out.v[i], r.base + full, sizeof(T));
    }
    return out;
}

template <class T>
void tcsg_emu_st(const tcsg_emu_rsrc &r, const u32 &off, const u32 &soff, const Lane<T> &v)
{
    for (int i = 0; i < tcsg_emu::W; ++i) {
        const uint64_t full = (uint64_t)off.v[i] + soff.v[i];
        if (full >> 32) { tcsg_emu_bad(); continue; }
        if (full + sizeof(T) > r.records) continue;  // out of range: dropped
        std::memcpy(r.base + full, &v.v[i], sizeof(T));
    }
}

// launch coordinates of the wave being emulated
struct tcsg_emu_dim3u { uint32_t x, y, z; };
struct tcsg_emu_dim3l { u32 x, y, z; };
extern tcsg_emu_dim3u blockIdx, blockDim, gridDim;
extern thread_local tcsg_emu_dim3l threadIdx;

// Workgroups of LDS-staging kernels: each wave runs on its own host thread and the waves
// meet at __syncthreads.  tcsg_emu_phase counts the barriers a wave has passed, so two
// accesses of one LDS element in the same phase by different waves, or a write and any
// other access in one phase, are a race the barriers do not order; such accesses, and
// ones outside the LDS or of elements never written, are counted as bad accesses.
extern thread_local unsigned long tcsg_emu_phase;

struct tcsg_emu_barrier_t {
    std::mutex m;
    std::condition_variable cv;
    unsigned n = 1, count = 0;
    unsigned long gen = 0;
    void start(unsigned waves) { std::lock_guard<std::mutex> l(m); n = waves; count = 0; }
};
inline tcsg_emu_barrier_t &tcsg_emu_bar()
{
    static tcsg_emu_barrier_t b;
    return b;
}
inline void __syncthreads()
{
    tcsg_emu_barrier_t &b = tcsg_emu_bar();
    std::unique_lock<std::mutex> l(b.m);
    const unsigned long g = b.gen;
    if (++b.count == b.n) {
        b.count = 0;
        ++b.gen;
        b.cv.notify_all();
    } else {
        b.cv.wait(l, [&] { return b.gen != g; });
    }
    ++tcsg_emu_phase;
}

// Runs fn(0..n-1) on n host threads with stacks large enough for the emulated waves'
// straight-line kernels (every per-lane value is a 64-lane array on the stack).
inline void tcsg_emu_run_waves(unsigned n, const std::function<void(unsigned)> &fn)
{
    struct Arg {
        const std::function<void(unsigned)> *fn;
        unsigned w;
    };
    std::vector<pthread_t> th(n);
    std::vector<Arg> args(n);
    pthread_attr_t at;
    pthread_attr_init(&at);
    pthread_attr_setstacksize(&at, (size_t)1 << 30);
    for (unsigned i = 0; i < n; ++i) {
        args[i] = Arg{&fn, i};
        if (pthread_create(&th[i], &at, [](void *p) -> void * {
                const Arg *a = (const Arg *)p;
                (*a->fn)(a->w);
                return nullptr;
            }, &args[i]) != 0)
            std::abort();
    }
    for (auto &t : th) pthread_join(t, nullptr);
    pthread_attr_destroy(&at);
}

template <class T>
struct tcsg_emu_lds {
    std::vector<T> v;
    // per element: phase + 1 of the last write and of the last read (0: none), and the
    // wave (its first lane's flat thread id) of the last write
    std::unique_ptr<std::atomic<unsigned long>[]> wph, rph, wwave;
    size_t n;
    explicit tcsg_emu_lds(size_t n_)
        : v(n_), wph(new std::atomic<unsigned long>[n_]), rph(new std::atomic<unsigned long>[n_]),
          wwave(new std::atomic<unsigned long>[n_]), n(n_) { reset(); }
    void reset()
    {
        for (size_t i = 0; i < n; ++i) { wph[i] = 0; rph[i] = 0; wwave[i] = 0; }
    }
};
inline unsigned long tcsg_emu_wave_id()
{
    return threadIdx.x.v[0] + blockDim.x * (threadIdx.y.v[0] + blockDim.y * threadIdx.z.v[0]);
}
template <class T>
Lane<T> tcsg_lds_ld(tcsg_emu_lds<T> &l, const u32 &idx)
{
    Lane<T> out(0);
    const unsigned long ph = tcsg_emu_phase + 1;
    for (int i = 0; i < tcsg_emu::W; ++i) {
        const uint32_t e = idx.v[i];
        if (e >= l.n) { tcsg_emu_bad(); continue; }
        const unsigned long wp = l.wph[e].load(std::memory_order_relaxed);
        if (wp == 0 || wp == ph) tcsg_emu_bad();   // never written, or written in this phase
        l.rph[e].store(ph, std::memory_order_relaxed);
        out.v[i] = l.v[e];
    }
    return out;
}
template <class T>
void tcsg_lds_st(tcsg_emu_lds<T> &l, const u32 &idx, const Lane<T> &val)
{
    const unsigned long ph = tcsg_emu_phase + 1, wave = tcsg_emu_wave_id();
    for (int i = 0; i < tcsg_emu::W; ++i) {
        const uint32_t e = idx.v[i];
        if (e >= l.n) { tcsg_emu_bad(); continue; }
        if (l.rph[e].load(std::memory_order_relaxed) == ph) tcsg_emu_bad();   // read in this phase
        if (l.wph[e].load(std::memory_order_relaxed) == ph && l.wwave[e].load(std::memory_order_relaxed) != wave)
            tcsg_emu_bad();                                                    // another wave wrote it
        l.wph[e].store(ph, std::memory_order_relaxed);
        l.wwave[e].store(wave, std::memory_order_relaxed);
        l.v[e] = val.v[i];
    }
}
