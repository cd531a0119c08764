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
#include <cmath>
#include <cstdint>
#include <cstring>
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

// A value the hardware holds in scalar registers (wave-uniform): lane 0's value; a lane
// that disagrees is counted in tcsg_emu_bad_access.
template <class T>
T tcsg_uni(const Lane<T> &v)
{
    for (int i = 1; i < tcsg_emu::W; ++i)
        if (v.v[i] != v.v[0]) ++tcsg_emu_bad_access;
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
        if (full >> 32) { ++tcsg_emu_bad_access; continue; }
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
        if (full >> 32) { ++tcsg_emu_bad_access; continue; }
        if (full + sizeof(T) > r.records) continue;  // out of range: dropped
        std::memcpy(r.base + full, &v.v[i], sizeof(T));
    }
}

// launch coordinates of the wave being emulated
struct tcsg_emu_dim3u { uint32_t x, y, z; };
struct tcsg_emu_dim3l { u32 x, y, z; };
extern tcsg_emu_dim3u blockIdx, blockDim, gridDim;
extern tcsg_emu_dim3l threadIdx;
