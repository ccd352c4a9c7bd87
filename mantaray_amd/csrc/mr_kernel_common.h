// mr_kernel_common.h -- pieces shared by the kernel translation units
// (mr_kernels.hip, mr_adjoint.hip): the per-lane view loader and the
// constants of the fast dispersion evaluation.  Include from .hip files only.
#pragma once

#include <type_traits>

#include "mr_device.h"

namespace mr {

// Word 3 of a raw buffer descriptor on the gfx9 family (gfx950 among them):
// DATA_FORMAT (bits 18:15) = 4, 32-bit, a valid format for untyped
// accesses; every other field zero.
constexpr int MR_BUFFER_DWORD3 = 0x00020000;
// cache-policy operand of the buffer intrinsics: bit 1 is the nontemporal
// (streaming) hint on gfx950
constexpr int MR_CPOL_NT = 2;

// s_waitcnt on the vector memory count alone: wait until at most N vector
// memory operations are outstanding (gfx9 encoding: vmcnt in bits 3:0,
// expcnt 6:4 and lgkmcnt 11:8 at their no-wait maxima, vmcnt bits 15:14
// zero).
template <int N>
__device__ __forceinline__ void vm_wait()
{
    static_assert(N >= 0 && N < 16, "vmcnt below 16");
    __builtin_amdgcn_s_waitcnt(0x0F70 | N);
}

// Copy a field view from device memory into per-lane registers.  The loads
// go through a lane-varying (always zero) offset so they are vector loads
// and the view lives in VGPRs for the whole kernel: the scalar register file
// cannot hold both views next to the loop's own scalars, and re-reading
// them with scalar loads puts a load latency into every evaluation.
template <class View>
__device__ __forceinline__ View load_view(const void *p)
{
    static_assert(sizeof(View) % 4 == 0, "view size");
    int zero;
    __asm__ volatile("v_mov_b32 %0, 0" : "=v"(zero));
    const int *src = reinterpret_cast<const int *>(p) + zero;
    int buf[sizeof(View) / 4];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(View) / 4); ++i) buf[i] = src[i];
    View v;
    __builtin_memcpy(&v, buf, sizeof(View));
    return v;
}

// Have a loaded view in registers at this point: each word is an input of
// an empty asm statement, so the compiler waits for the view's loads here
// and knows them complete in the code that follows (a loop after this does
// not wait for them again, nor for whatever else is in flight then).
template <class View>
__device__ __forceinline__ void settle(const View &v)
{
    int buf[sizeof(View) / 4];
    __builtin_memcpy(buf, &v, sizeof(View));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(View) / 4); ++i) __asm__ volatile("" ::"v"(buf[i]));
}

static __device__ double mr_fast_consts[sizeof(FastConsts) / sizeof(double)] = {
    1.4426950408889634,                    // 1/ln 2
    6.93147180369123816490e-01,            // ln 2 split, upper part
    1.90821492927058770002e-10,            // ln 2 - upper part
    1400.0,
    // expm1(f)/f interpolated at the 10 Chebyshev nodes of [-ln2/2, ln2/2]
    // (computed in 50-digit arithmetic and rounded), c_9 .. c_0
    2.7620075879983367e-07,
    2.7632640675430236e-06,
    2.4801521322368692e-05,
    1.984119064754424e-04,
    1.3888888917196719e-03,
    8.333333367311603e-03,
    4.1666666666624164e-02,
    1.6666666666615648e-01,
    5.000000000000001e-01,
    1.0000000000000013,
    G,
    0.5 * G,
    -2.0 * G,
    __builtin_nan(""),
};

// t[0] = t0, t[s] = t[s-1] + dt for s <= n
__device__ __forceinline__ void time_column(double t0, double dt, int64_t n, double *t)
{
    double v = t0;
    t[0] = v;
    for (int64_t s = 1; s <= n; ++s) {
        v = v + dt;
        t[s] = v;
    }
}

// Field kinds and math mode as compile-time parameters.
template <int BK, int CK, int MATH>
struct Kinds {
    static constexpr int bk = BK, ck = CK, math = MATH;
};

// Call f(Kinds<BK, CK, MATH>{}) for the run-time field kinds and math mode.
template <class F>
void dispatch_kinds(int bk, int ck, int math, F &&f)
{
    auto by_math = [&](auto b, auto c) {
        constexpr int BK = decltype(b)::value, CK = decltype(c)::value;
        if (math == MATH_FAST) f(Kinds<BK, CK, MATH_FAST>{});
        else f(Kinds<BK, CK, MATH_LITERAL>{});
    };
    auto by_current = [&](auto b) {
        if (ck == CK_CONSTANT) by_math(b, std::integral_constant<int, CK_CONSTANT>{});
        else by_math(b, std::integral_constant<int, CK_GRID>{});
    };
    switch (bk) {
    case BK_CONSTANT: by_current(std::integral_constant<int, BK_CONSTANT>{}); break;
    case BK_SLOPE: by_current(std::integral_constant<int, BK_SLOPE>{}); break;
    case BK_ARRAY: by_current(std::integral_constant<int, BK_ARRAY>{}); break;
    default: by_current(std::integral_constant<int, BK_GRID>{}); break;
    }
}

}  // namespace mr
