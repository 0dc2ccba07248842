// CPU emulation of the HIP device subset the engine's kernels use, for
// checking kernel logic on the host (sanitizers, no GPU).  One block runs at
// a time, one std::thread per lane; __syncthreads is a block barrier and
// every cross-lane operation a barrier of the lane's wavefront, so a shuffle
// reached by only part of a wavefront deadlocks here instead of silently
// misbehaving.  Static __shared__ arrays become function statics (one block
// at a time).
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#define GEODIC_SIMT 1
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define GEODIC_BATCH_ATTRS(W, OCC)
#define GEODIC_PLAIN_ATTRS
#define GEODIC_SCHED_BARRIER()
#define GEODIC_DYN_LDS(T, name) T* name = (T*)::simt::g_dyn_lds

using std::max;
using std::min;

struct alignas(16) float4 {
    float x, y, z, w;
};
struct alignas(8) float2 {
    float x, y;
};
inline float2 make_float2(float x, float y) { return {x, y}; }
inline float __int_as_float(int v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}
inline int __float_as_int(float f) {
    int v;
    std::memcpy(&v, &f, 4);
    return v;
}
struct double2 {
    double x, y;
};
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline double2 make_double2(double x, double y) { return {x, y}; }

namespace simt {
constexpr int kLanes = 64;
constexpr int kMaxThreads = 1024;
inline std::barrier<>* g_bar = nullptr;
inline std::barrier<>* g_wave_bar[kMaxThreads / kLanes] = {};
inline void* g_dyn_lds = nullptr;
inline uint64_t g_xchg[kMaxThreads];
struct Idx {
    int x = 0, y = 0, z = 0;
};
inline thread_local Idx t_thread;
inline Idx g_block, g_blockdim{kLanes, 1, 1}, g_griddim{1, 1, 1};

inline void sync() { g_bar->arrive_and_wait(); }
inline void wave_sync() { g_wave_bar[t_thread.x / kLanes]->arrive_and_wait(); }

// Cross-lane read within the calling thread's wavefront (every lane of the
// wavefront must reach it: it is a wavefront barrier here).
template <class T>
T xchg(T v, int src) {
    static_assert(sizeof(T) <= 8, "exchange slot");
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    g_xchg[t_thread.x] = bits;
    wave_sync();
    uint64_t o = g_xchg[(t_thread.x & ~(kLanes - 1)) | (src & (kLanes - 1))];
    wave_sync();
    T r;
    std::memcpy(&r, &o, sizeof(T));
    return r;
}

// Runs `body` for blocks 0..grid-1, one block (`threads` lanes) at a time.
inline void launch(int grid, int threads, size_t dyn_lds_bytes, const std::function<void()>& body) {
    // 16-byte aligned, exactly the requested size (ASan sees overruns)
    unsigned char* lds = static_cast<unsigned char*>(::operator new(dyn_lds_bytes + 16, std::align_val_t(16)));
    std::memset(lds, 0xCD, dyn_lds_bytes);
    g_dyn_lds = lds;
    g_griddim.x = grid;
    g_blockdim.x = threads;
    for (int b = 0; b < grid; ++b) {
        g_block.x = b;
        std::barrier<> bar(threads);
        g_bar = &bar;
        std::vector<std::unique_ptr<std::barrier<>>> wb;
        for (int w = 0; w < (threads + kLanes - 1) / kLanes; ++w) {
            wb.emplace_back(new std::barrier<>(std::min(kLanes, threads - w * kLanes)));
            g_wave_bar[w] = wb.back().get();
        }
        std::vector<std::thread> th;
        for (int l = 0; l < threads; ++l)
            th.emplace_back([l, &body] {
                t_thread.x = l;
                body();
            });
        for (auto& t : th) t.join();
    }
    ::operator delete(lds, std::align_val_t(16));
}
}  // namespace simt

#define threadIdx (::simt::t_thread)
#define blockIdx (::simt::g_block)
#define blockDim (::simt::g_blockdim)
#define gridDim (::simt::g_griddim)

inline void __syncthreads() { simt::sync(); }
inline void __builtin_amdgcn_wave_barrier() { simt::wave_sync(); }
// memory fence (scope ignored: every lane is a thread of this process)
inline void __builtin_amdgcn_fence(int, const char*) { std::atomic_thread_fence(std::memory_order_seq_cst); }
template <class T>
T __shfl_xor(T v, int off, int = 64) {
    return simt::xchg(v, simt::t_thread.x ^ off);
}
template <class T>
T __shfl(T v, int src, int = 64) {
    return simt::xchg(v, src);
}
inline int __builtin_amdgcn_readfirstlane(int v) { return simt::xchg(v, 0); }
inline int __builtin_amdgcn_readlane(int v, int lane) { return simt::xchg(v, lane); }

// DPP move on gfx9 encodings used by the kernels: quad_perm (0x00-0xff),
// row_shl:n (0x101-0x10f), row_shr:n (0x111-0x11f), row_ror:n (0x121-0x12f),
// row_bcast:15 (0x142), row_bcast:31 (0x143).  Lanes of rows outside
// row_mask, or of 4-lane banks outside bank_mask, keep `old`; a lane whose
// source is outside its row reads 0 with bound_ctrl, else keeps old.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    const int lane = simt::t_thread.x & 63;
    const int row = lane >> 4, rl = lane & 15;
    int from = -1;
    if (ctrl < 0x100) {
        from = (lane & ~3) | ((ctrl >> (2 * (lane & 3))) & 3);
    } else if (ctrl >= 0x101 && ctrl <= 0x10f) {
        const int n = ctrl & 15;
        from = (rl + n < 16) ? lane + n : -1;
    } else if (ctrl >= 0x111 && ctrl <= 0x11f) {
        const int n = ctrl & 15;
        from = (rl >= n) ? lane - n : -1;
    } else if (ctrl >= 0x121 && ctrl <= 0x12f) {
        const int n = ctrl & 15;
        from = 16 * row + ((rl - n + 16) & 15);
    } else if (ctrl == 0x142) {
        from = row >= 1 ? 16 * row - 1 : -1;
    } else if (ctrl == 0x143) {
        from = row >= 2 ? 31 : -1;
    } else {
        std::abort();
    }
    const int v = simt::xchg(src, from < 0 ? lane : from);  // xchg stays within the wavefront
    if (!((row_mask >> row) & 1)) return old;
    if (!((bank_mask >> (rl >> 2)) & 1)) return old;
    if (from < 0) return bound_ctrl ? 0 : old;
    return v;
}

// v_permlane32_swap / v_permlane16_swap: element 0 is the new VDST, element
// 1 the new VSRC0.  32: VDST lanes 32-63 <-> VSRC0 lanes 0-31.  16: VDST's
// odd 16-lane rows <-> VSRC0's even rows.
struct simt_u32x2 {
    unsigned v[2];
    unsigned operator[](int i) const { return v[i]; }
};
inline simt_u32x2 __builtin_amdgcn_permlane32_swap(unsigned dst, unsigned src, bool, bool) {
    const int lane = simt::t_thread.x & 63;
    const unsigned d_peer = simt::xchg(dst, lane ^ 32);
    const unsigned s_peer = simt::xchg(src, lane ^ 32);
    simt_u32x2 r;
    r.v[0] = lane < 32 ? dst : s_peer;  // upper VDST lanes take VSRC0's lower lanes
    r.v[1] = lane < 32 ? d_peer : src;  // lower VSRC0 lanes take VDST's upper lanes
    return r;
}
inline simt_u32x2 __builtin_amdgcn_permlane16_swap(unsigned dst, unsigned src, bool, bool) {
    const int lane = simt::t_thread.x & 63;
    const unsigned d_peer = simt::xchg(dst, lane ^ 16);
    const unsigned s_peer = simt::xchg(src, lane ^ 16);
    const bool odd = (lane >> 4) & 1;
    simt_u32x2 r;
    r.v[0] = odd ? s_peer : dst;   // odd VDST rows take VSRC0's even rows
    r.v[1] = odd ? src : d_peer;   // even VSRC0 rows take VDST's odd rows
    return r;
}
// v_rcp_f64 / v_rsq_f64 estimates: the exact quotient and root here (the
// device refines its estimates by Newton steps to within an ulp or two)
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
inline double __builtin_inf() { return INFINITY; }
// system-scope release store (a list solve's done flag)
#define __HIP_MEMORY_SCOPE_SYSTEM 0
template <class T>
inline void __hip_atomic_store(T* p, T v, int, int) {
    __atomic_store_n(p, v, __ATOMIC_RELEASE);
}
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline double __longlong_as_double(long long v) {
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}
inline int __double2loint(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return (int)(uint32_t)b;
}
inline int __double2hiint(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return (int)(uint32_t)(b >> 32);
}
inline double __hiloint2double(int hi, int lo) {
    uint64_t b = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
