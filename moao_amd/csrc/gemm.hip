// MFMA GEMM for gfx950 (FP64 via v_mfma_f64_16x16x4f64, FP32 via v_mfma_f32_16x16x4f32).
//
// Workgroup: 256 lanes = 4 waves in a 2x2 arrangement over a 128x128 C tile; each wave owns a
// 64x64 sub-tile = 4x4 MFMA blocks of 16x16 held in accumulator registers.  K advances in
// slices of 8.  MFMA operand layout (checked on the device): lane l supplies A[l%16][k0+l/16]
// and B[k0+l/16][l%16]; accumulator item t of lane l is C[r][l%16] with r = 4*(l/16)+t for f32
// and r = l/16 + 4t for f64.
//
// Two kernels share that structure:
//  * gemm_dl_kernel (every launch with op(B) = B or B^T, and symmetric B when the caller
//    supplies full copies of its diagonal blocks): buffer loads write LDS directly
//    (buffer_load_dword ... lds), so slices need no register staging; kStages slices are in
//    flight and each slice costs one vmcnt wait and one plain s_barrier.  The buffer
//    descriptors' range checks zero-fill what lies past the matrix (the range includes the
//    scalar offset on gfx950, so every column gets a descriptor of its own).  Measured
//    16384^3 FP64: 74 TF/s (94% of the MFMA peak).
//  * gemm_kernel (fallback: symmetric B without diagonal copies, K x N slabs whose byte
//    offsets would overflow 32 bits): register-staged, double-buffered LDS, the next slice's
//    global loads issued before the current slice's MFMAs.
// A launch with fewer than two 128-row tiles per CU (panel solves of the leaves, narrow
// updates) uses 64-row tiles (each wave 32x64, up to 4 workgroups per CU), which doubles the
// workgroups that fill the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "gemm.hpp"
#include "profile.hpp"

namespace moao {
namespace {

constexpr int BM = 128, BN = 128, BK = 8;
constexpr int PAD = 2;
constexpr int NT = 256;
#ifndef MOAO_GEMM_STAGES
#define MOAO_GEMM_STAGES 3
#endif
constexpr int kStages = MOAO_GEMM_STAGES;  // direct-to-LDS pipeline depth (slices in flight + 1)

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator item t of lane l holds row (l/16) + 4t of the 16x16 block
    static __device__ __forceinline__ int row(int lane, int t) { return (lane >> 4) + 4 * t; }
};

template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator item t of lane l holds row 4*(l/16) + t of the 16x16 block
    static __device__ __forceinline__ int row(int lane, int t) { return 4 * (lane >> 4) + t; }
};

// TB: 0 = B is K x N, 1 = B is N x K (op = transpose), 2 = B symmetric K x K, lower stored.
// Loads are branch-free: addresses are clamped into the matrix and out-of-range values zeroed
// by a select, so the compiler issues every load of a slice back to back.
// WMB: 16-row MFMA blocks per wave along M (4: 128-row tile; 2: 64-row tile for launches
// too small to fill the device with 128-row tiles).
template <typename T, int TB, int WMB>
__global__ void __launch_bounds__(NT, WMB == 4 ? 2 : 4) gemm_kernel(GemmArgs<T> g) {
    using acc_t = typename Mfma<T>::acc_t;
    constexpr int TM = 32 * WMB;  // rows of the C tile
    __shared__ T As[2][BK][TM + PAD];
    __shared__ T Bs[2][BK][BN + PAD];

    if (g.skip_if && *(volatile const int*)g.skip_if != 0) return;
    const GemmSeg<T> s = g.seg[blockIdx.z];
    const int m0 = blockIdx.x * TM;
    const int n0 = blockIdx.y * BN;
    if (m0 >= s.M) return;
    const bool tri = g.tri && (blockIdx.z == 0 || g.tri_all);
    if (tri && m0 + TM - 1 + g.tri_off < n0) return;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int K = g.K, N = g.N, M = s.M;
    const int Mc = M - 1, Nc = N - 1, Kc = K > 0 ? K - 1 : 0;

    acc_t acc[WMB][4];
#pragma unroll
    for (int i = 0; i < WMB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // Global -> register staging.  A (and B^T): a wave runs down 64 contiguous rows of one
    // column; which column is wave-uniform (readfirstlane), so the column base is a scalar and
    // the per-lane part is a row offset computed once per tile: no per-load address VALU.
    // Raw loads go to registers with a validity mask applied at the LDS store (after the
    // slice's MFMAs), so no wait sits between the loads.
    constexpr int RV = BK / 2, CP = NT / BK;  // B values per thread, columns per pass
    constexpr int KA = NT / TM, RA = BK / KA;  // A: k columns per pass, values per thread
    T ra[RA], rb[RV];
    unsigned ma = 0, mb = 0;
    const int lrow = tid & 127;
    const int lk = __builtin_amdgcn_readfirstlane(tid >> 7);  // 0 for waves 0-1, 1 for 2-3
    const int arl = tid % TM;
    const int akl = __builtin_amdgcn_readfirstlane(tid / TM);
    const int bk = tid % BK, bn = tid / BK;                   // B (K x N) pattern
    const long lda = s.lda, ldb = g.ldb;
    const int arow = m0 + arl;
    const unsigned aoff = (unsigned)min(arow, Mc) * (unsigned)sizeof(T);  // bytes
    const bool arow_ok = arow < M;
    const T* __restrict__ A = s.A;
    const T* __restrict__ Bm = s.B ? s.B : g.B;
    // TB=0 / crossing-slice pattern: per-lane column offsets, computed once per tile
    long bcol[RV];
#pragma unroll
    for (int r = 0; r < RV; ++r) bcol[r] = (long)min(n0 + bn + CP * r, Nc) * ldb;
    const unsigned bcoloff = (unsigned)min(n0 + lrow, Nc) * (unsigned)sizeof(T);  // bytes
    const bool bcol_ok = n0 + lrow < N;

    int bpat = TB == 1 ? 1 : 0;  // layout of the slice held in rb
    auto load = [&](int k0) {
        ma = mb = 0;
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            const int k = k0 + akl + KA * r;
            const T* col = A + (long)min(k, Kc) * lda;  // wave-uniform
            ra[r] = *(const T*)((const char*)col + aoff);
            ma |= (unsigned)(arow_ok && k < K) << r;
        }
        int pat = TB == 1 ? 1 : 0;
        if (TB == 2) {
            if (k0 + BK - 1 < n0) pat = 1;            // slice wholly above the diagonal
            else if (k0 >= n0 + BN - 1) pat = 0;      // wholly below
            else pat = 2;                             // crossing: per-element triangle
        }
        bpat = pat == 1 ? 1 : 0;
        if (pat == 1) {
#pragma unroll
            for (int r = 0; r < RV; ++r) {
                const int k = k0 + lk + 2 * r;
                const T* col = Bm + (long)min(k, Kc) * ldb;  // wave-uniform
                rb[r] = *(const T*)((const char*)col + bcoloff);
                mb |= (unsigned)(bcol_ok && k < K) << r;
            }
        } else if (pat == 0) {
            const int k = k0 + bk, kc = min(k, Kc);
#pragma unroll
            for (int r = 0; r < RV; ++r) {
                rb[r] = Bm[bcol[r] + kc];
                mb |= (unsigned)(n0 + bn + CP * r < N && k < K) << r;
            }
        } else {
            const int k = k0 + bk, kc = min(k, Kc);
#pragma unroll
            for (int r = 0; r < RV; ++r) {
                const int cc = min(n0 + bn + CP * r, Nc);
                rb[r] = kc >= cc ? Bm[bcol[r] + kc] : Bm[cc + (long)kc * ldb];
                mb |= (unsigned)(n0 + bn + CP * r < N && k < K) << r;
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int r = 0; r < RA; ++r) As[buf][akl + KA * r][arl] = (ma >> r & 1) ? ra[r] : T(0);
        if (bpat) {
#pragma unroll
            for (int r = 0; r < RV; ++r)
                Bs[buf][lk + 2 * r][lrow] = (mb >> r & 1) ? rb[r] : T(0);
        } else {
#pragma unroll
            for (int r = 0; r < RV; ++r)
                Bs[buf][bk][bn + CP * r] = (mb >> r & 1) ? rb[r] : T(0);
        }
    };

    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    int cur = 0;
    const int ar = wm * 16 * WMB + (lane & 15);
    const int bc = wn * 64 + (lane & 15);
    const int kq = lane >> 4;
    for (int kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            const int k = kk * 4 + kq;
            T a[WMB], b[4];
#pragma unroll
            for (int i = 0; i < WMB; ++i) a[i] = As[cur][k][ar + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[cur][k][bc + 16 * j];
#pragma unroll
            for (int i = 0; i < WMB; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::mac(a[i], b[j], acc[i][j]);
        }
        if (kt + 1 < nk) store(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }

    // Epilogue: C = alpha*acc + beta*C (C not read when beta == 0).
    const T alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < WMB; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn * 64 + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = m0 + wm * 16 * WMB + i * 16 + Mfma<T>::row(lane, t);
                if (row >= M) continue;
                if (tri && row + g.tri_off < col) continue;
                T* c = s.C + row + (long)col * s.ldc;
                T v = alpha * acc[i][j][t];
                if (beta != T(0)) v += beta * *c;
                *c = v;
            }
        }
    }
}

// Direct-to-LDS kernel: S slices in flight.  A and op(B) = B^T tiles are column panels of 128
// (A: TM) rows: one descriptor per column, so the descriptor's range check zero-fills rows past
// the matrix.  op(B) = B (K x N) tiles are BK-row slabs of 128 columns: one descriptor per slab
// whose range ends at the slab's last element, 4 (FP64) or 8 (FP32) columns per 256-byte load,
// column groups padded in LDS.  Symmetric B switches between the two per slice.
// The LDS pointer passes through an empty asm so the compiler cannot trace it to one LDS array:
// with that provenance it would put a vmcnt(0) before every later read of the same array and
// serialize the pipeline.  Ordering is explicit instead: wait_vm + bar before each slice's reads.
template <typename T>
__device__ __forceinline__ void dl_load(__amdgpu_buffer_rsrc_t r, T* lds, int voff) {
    __attribute__((address_space(3))) void* l = (__attribute__((address_space(3))) void*)lds;
    asm("" : "+s"(l));  // the 32-bit LDS address, so no generic-to-LDS conversion follows
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 4, voff, 0, 0, 0);
}
// s_waitcnt vmcnt(VM) alone (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4 and lgkmcnt
// 11:8 left at their maxima).
template <int VM>
__device__ __forceinline__ void wait_vm() {
    static_assert(VM >= 0 && VM < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((VM & 15) | ((VM >> 4) << 14) | (7 << 4) | (15 << 8));
}
// Workgroup barrier without the memory fence of __syncthreads (which would also drain vmcnt);
// the empty asm statements keep the compiler from moving LDS accesses across it.
__device__ __forceinline__ void bar() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Tile order of a launch.  The dispatcher hands workgroups to the 8 XCDs round robin by launch
// index (index % 8), and each XCD has an L2 of its own.  The launch index L is re-mapped so
// that the k-th run of kXcdChunk workgroups an XCD receives is kXcdChunk consecutive tiles of
// a grouped order (strips of kGroupM tile rows, walked column by column): about
// kGroupM x (kXcdChunk / kGroupM) tiles whose A row panels and B column panels the XCD's L2
// shares, while the 8 XCDs still sweep the grid together (same balance as the plain order).
// The map is a bijection on [0, tiles): launch indices past the last whole round of
// 8 * kXcdChunk keep their place.
#ifndef MOAO_GEMM_REMAP
#define MOAO_GEMM_REMAP 1
#endif
#ifndef MOAO_GEMM_XCD_CHUNK
#define MOAO_GEMM_XCD_CHUNK 32
#endif
#ifndef MOAO_GEMM_GROUP_M
#define MOAO_GEMM_GROUP_M 8
#endif
constexpr int kXcds = 8;
constexpr int kXcdChunk = MOAO_GEMM_XCD_CHUNK;
constexpr int kGroupM = MOAO_GEMM_GROUP_M;

struct TileIdx {
    int x, y, z;
};

__host__ __device__ inline TileIdx tile_order(long L, int gx, int gy, int gz) {
    const long total = (long)gx * gy * gz;
    if (kXcdChunk > 0) {
        const long round = (long)kXcds * kXcdChunk;
        if (L < total / round * round) {
            const long c = L % kXcds, q = L / kXcds;
            L = (q / kXcdChunk) * round + c * kXcdChunk + q % kXcdChunk;
        }
    }
    const long slice = (long)gx * gy;
    const int z = (int)(L / slice);
    const long r = L % slice;
    const long strip = (long)kGroupM * gy;
    const int x0 = (int)(r / strip) * kGroupM;
    const int gm = kGroupM < gx - x0 ? kGroupM : gx - x0;
    const long rr = r - (long)x0 * gy;
    return TileIdx{x0 + (int)(rr % gm), (int)(rr / gm), z};
}

__device__ __forceinline__ TileIdx block_tile() {
    if (!MOAO_GEMM_REMAP) return TileIdx{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
    const long L = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    return tile_order(L, gridDim.x, gridDim.y, gridDim.z);
}

template <typename T, int TB, int S, int TM>
__global__ void __launch_bounds__(NT, TM == 128 ? 2 : 4) gemm_dl_kernel(GemmArgs<T> g) {
    using acc_t = typename Mfma<T>::acc_t;
    constexpr int WMB = TM / 32;                       // 16-row MFMA blocks per wave along M
    constexpr int LDA = TM + PAD, LDB = BN + PAD;
    constexpr int QA = TM * (int)sizeof(T) / 256;      // 256-byte pieces per A tile column
    constexpr int QB = BN * (int)sizeof(T) / 256;      // ... per B^T tile column
    constexpr int CPI = 256 / (BK * (int)sizeof(T));   // K x N: columns per 256-byte load
    constexpr int LPC = 64 / CPI;                      // K x N: lanes per column
    constexpr int GS = CPI * BK + 16 / (int)sizeof(T);  // K x N: LDS elements per column group
    constexpr int NG = BN / CPI;                       // K x N: column groups per tile
    constexpr int IPWA = BK / 2 * QA;                  // loads per A wave per slice
    constexpr int IPWB = TB == 0 ? NG / 2 : BK / 2 * QB;  // loads per B wave per slice
    static_assert(TB != 2 || BK / 2 * QB == NG / 2, "both B patterns issue the same loads");
    constexpr int BSZ = TB == 1 ? BK * LDB : (TB == 0 ? NG * GS : std::max(BK * LDB, NG * GS));
    __shared__ T As[S][BK][LDA];
    __shared__ T Bs[S][BSZ];

    if (g.skip_if && *(volatile const int*)g.skip_if != 0) return;
    const TileIdx tl = block_tile();
    const GemmSeg<T> s = g.seg[tl.z];
    const int m0 = tl.x * TM;
    const int n0 = tl.y * BN;
    if (m0 >= s.M) return;
    const bool tri = g.tri && (tl.z == 0 || g.tri_all);
    if (tri && m0 + TM - 1 + g.tri_off < n0) return;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const int K = g.K, N = g.N, M = s.M;

    acc_t acc[WMB][4];
#pragma unroll
    for (int i = 0; i < WMB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // waves 0-1 load A (half the slice's columns each), waves 2-3 load B (half a slice each).
    // B slices come as column panels (op(B) = B^T, and symmetric B above the tile's diagonal
    // block: column k of the lower triangle, rows n0..) or as row slabs (op(B) = B, and
    // symmetric B below the diagonal block: rows k of columns n0..); symmetric slices that
    // cross the diagonal block read its full symmetric copy in g.symdiag.
    const bool isA = wave < 2;
    const T* Bm = s.B ? s.B : g.B;
    const T* base = isA ? s.A + m0 : (TB == 0 ? Bm + (long)n0 * g.ldb : Bm + n0);
    const long ld = isA ? s.lda : g.ldb;
    const T* diagblk = TB == 2 ? g.symdiag + (long)(n0 / BN) * BN * BN : nullptr;
    const int nv = isA ? min(M - m0, TM) : min(N - n0, BN);  // valid rows (cols for K x N)
    const int half = wave & 1;
    // K x N: lane -> (column lane/LPC of the group, dword lane%LPC of that column's slab)
    const int vkn = (lane / LPC) * (int)(ld * (long)sizeof(T)) + (lane % LPC) * 4;
    const int gstride = CPI * (int)(ld * (long)sizeof(T));  // bytes between column groups

    // column panels: one descriptor per column (column k of src sits at src + (k-kofs)*lds),
    // Q pieces of 256 bytes each
    auto colpanel = [&](auto qc, T* dst, int ldl, const T* src, long lds, int kofs, int k0) {
        constexpr int Q = decltype(qc)::value;
#pragma unroll
        for (int c = 0; c < BK / 2; ++c) {
            const int kk = half * (BK / 2) + c;
            const int kcl = min(k0 + kk, K - 1);  // clamp past K (masked at use)
            const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(src + (long)(kcl - kofs) * lds), 0, nv * (int)sizeof(T), 0x00020000);
#pragma unroll
            for (int q = 0; q < Q; ++q)
                dl_load(r, (T*)((char*)(dst + kk * ldl) + q * 256), lane * 4 + q * 256);
        }
    };
    const int nk = (K + BK - 1) / BK;
    auto issue = [&](int kt) {
        const int k0 = kt * BK;
        const bool below = TB == 2 && k0 >= n0 + BN;
        const bool cross = TB == 2 && !below && k0 + BK > n0;
        if (isA) {
            colpanel(std::integral_constant<int, QA>{}, &As[kt % S][0][0], LDA, base, ld, 0, k0);
        } else if (TB == 1 || (TB == 2 && !below)) {
            colpanel(std::integral_constant<int, QB>{}, &Bs[kt % S][0], LDB,
                     cross ? diagblk : base, cross ? (long)BN : ld, cross ? n0 : 0, k0);
        } else {
            // slab rows k0..k0+BK-1 of columns n0..n0+nv-1; the range ends at the last valid
            // element so nothing past the matrix is read
            const T* slab = TB == 0 ? base : Bm + (long)n0 * ld;
            const int kv = min(BK, K - k0);
            const int nrec = ((nv - 1) * (int)ld + kv) * (int)sizeof(T);
            const __amdgpu_buffer_rsrc_t r =
                __builtin_amdgcn_make_buffer_rsrc((void*)(slab + k0), 0, nrec, 0x00020000);
            T* dst = &Bs[kt % S][0];
#pragma unroll
            for (int j = 0; j < IPWB; ++j) {
                const int grp = half * IPWB + j;
                dl_load(r, dst + grp * GS, vkn + grp * gstride);
            }
        }
    };

    for (int p = 0; p < S - 1; ++p)
        if (p < nk) issue(p);

    const int ar = wm * 16 * WMB + (lane & 15);
    const int kq = lane >> 4;
    // B operand element (k, n = wn*64 + 16j + lane%16) in LDS
    const int bn = wn * 64 + (lane & 15);
    const int boffC = bn, boffS = (bn / CPI) * GS + (bn % CPI) * BK;  // panel / slab layouts
    constexpr int BJS = (16 / CPI) * GS;                // slab layout: step per j
    for (int kt = 0; kt < nk; ++kt) {
        // this wave's loads of slice kt are done once at most the later slices' are pending
        if (kt + S - 2 >= nk) wait_vm<0>();
        else if (isA) wait_vm<(S - 2) * IPWA>();
        else wait_vm<(S - 2) * IPWB>();
        bar();
        if (kt + S - 1 < nk) issue(kt + S - 1);
        const int cur = kt % S;
        const bool tail = kt * BK + BK > K;
        const bool slab = TB == 0 || (TB == 2 && kt * BK >= n0 + BN);
        const int boff = slab ? boffS : boffC, bj = slab ? BJS : 16, bks = slab ? 1 : LDB;
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            const int k = kk * 4 + kq;
            T a[WMB], b[4];
#pragma unroll
            for (int i = 0; i < WMB; ++i) a[i] = As[cur][k][ar + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[cur][boff + bj * j + bks * k];
            if (tail) {
                const bool ok = kt * BK + k < K;
#pragma unroll
                for (int i = 0; i < WMB; ++i) a[i] = ok ? a[i] : T(0);
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = ok ? b[j] : T(0);
            }
#pragma unroll
            for (int i = 0; i < WMB; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::mac(a[i], b[j], acc[i][j]);
        }
    }

    const T alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < WMB; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn * 64 + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = m0 + wm * 16 * WMB + i * 16 + Mfma<T>::row(lane, t);
                if (row >= M) continue;
                if (tri && row + g.tri_off < col) continue;
                T* c = s.C + row + (long)col * s.ldc;
                T v = alpha * acc[i][j][t];
                if (beta != T(0)) v += beta * *c;
                *c = v;
            }
        }
    }
}

// Split-BF16 products.  Each FP32 operand element a is split as hi = bf16(a) and
// lo = bf16(a - hi) (round to nearest even, v_cvt_pk_bf16_f32), and the product accumulates
// lo*hi + hi*lo + hi*hi in FP32 on v_mfma_f32_16x16x32_bf16: about 16 significant bits per
// operand (the dropped lo*lo term is 2^-16 relative) at three BF16 MFMAs per 16x16x32 step,
// which together take 3/8 of the time of the eight FP32 MFMAs the same step costs.  The MFMA
// pairs A's item i of lane l with B's item i of the lane of the same l/16, so any common k
// order within a chunk gives the same sum; the output layout is that of the FP32 16x16 MFMA
// (row 4*(l/16)+t, column l%16).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int KC = 32;  // k per chunk (one 16x16x32 step)

// Split-BF16 kernel on pre-split planes.  Each stage holds 32 k of the A tile (TM rows) and of
// the B tile (128 rows of op(B)^T), hi and lo planes, in 16-row blocks of 1 KB laid out exactly
// as the 16x16x32 MFMA takes its operands: lane l of a fragment read supplies row l % 16 and
// k = 8 (l / 16) .. + 7, i.e. bytes [16 l, 16 l + 16) of the block, so every fragment is one
// conflict-free 16-byte LDS read.  The blocks are filled by 16-byte buffer loads straight into
// LDS (buffer_load_dwordx4 ... lds, one 1 KB block per instruction, no registers), kS2 stages
// deep.  Products accumulate in FP32 in the MFMA; every kFlush stages the accumulators are
// added into a second set of FP32 sums and restarted, so no accumulation chain runs longer
// than 256 k.
constexpr int kS2 = 3;      // stages of the planes kernel
constexpr int kFlush = 8;   // stages between flushes of the MFMA accumulators

__device__ __forceinline__ void dl_load16(__amdgpu_buffer_rsrc_t r, void* lds, int voff) {
    __attribute__((address_space(3))) void* l = (__attribute__((address_space(3))) void*)lds;
    asm("" : "+s"(l));
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 16, voff, 0, 0, 0);
}

template <int TM>
__global__ void __launch_bounds__(NT, TM == 128 ? 1 : 2) gemm_split2_kernel(GemmArgs<float> g) {
    constexpr int WMB = TM / 32;          // 16-row MFMA blocks per wave along M
    constexpr int ABLK = TM / 16;         // 16-row blocks of the A tile
    constexpr int NBLK = ABLK + BN / 16;  // blocks per plane per stage
    constexpr int IPW = 2 * NBLK / 4;     // 16-byte DMA instructions per wave per stage
    __shared__ __attribute__((aligned(16))) unsigned char lds[kS2][2][NBLK][1024];

    if (g.skip_if && *(volatile const int*)g.skip_if != 0) return;
    const GemmSeg<float> s = g.seg[blockIdx.z];
    const int m0 = blockIdx.x * TM;
    const int n0 = blockIdx.y * BN;
    if (m0 >= s.M) return;
    const bool tri = g.tri && (blockIdx.z == 0 || g.tri_all);
    if (tri && m0 + TM - 1 + g.tri_off < n0) return;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const int K = g.K, N = g.N, M = s.M;

    f32x4 acc[WMB][4], sum[WMB][4];
#pragma unroll
    for (int i = 0; i < WMB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = sum[i][j] = f32x4{0, 0, 0, 0};

    // DMA instruction q of a stage: plane q / NBLK, block q % NBLK (A blocks, then B blocks).
    // Lane l loads row (block row + l % 16), clamped into the operand, 16 bytes at k group l / 16.
    const int lr = lane & 15, lkg = lane >> 4;
    auto issue = [&](int kt) {
        const int k0 = kt * KC;
#pragma unroll
        for (int u = 0; u < IPW; ++u) {
            const int q = wave + 4 * u;
            const int pl = q / NBLK, blk = q % NBLK;
            const bool isA = blk < ABLK;
            const Planes& P = isA ? s.ap : g.bp;
            const int rows = isA ? M : N;
            const int rb = min(isA ? m0 + 16 * blk : n0 + 16 * (blk - ABLK), rows - 1);
            const int rl = min(rb + lr, rows - 1) - rb;
            const unsigned short* base = P.hi + (pl ? P.ps : 0) + (long)rb * P.ld + k0;
            const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
                (void*)base, 0, (int)(16 * P.ld * 2), 0x00020000);
            dl_load16(r, &lds[kt % kS2][pl][blk][0], (rl * (int)P.ld + 8 * lkg) * 2);
        }
    };

    const int nk = (K + KC - 1) / KC;
    for (int p = 0; p < kS2 - 1; ++p)
        if (p < nk) issue(p);
    for (int kt = 0; kt < nk; ++kt) {
        if (kt + kS2 - 2 >= nk) wait_vm<0>();
        else wait_vm<(kS2 - 2) * IPW>();
        bar();
        if (kt + kS2 - 1 < nk) issue(kt + kS2 - 1);
        const int cur = kt % kS2;
        bf16x8 ah[WMB], al[WMB], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < WMB; ++i) {
            ah[i] = *(const bf16x8*)&lds[cur][0][wm * WMB + i][16 * lane];
            al[i] = *(const bf16x8*)&lds[cur][1][wm * WMB + i][16 * lane];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bh[j] = *(const bf16x8*)&lds[cur][0][ABLK + wn * 4 + j][16 * lane];
            bl[j] = *(const bf16x8*)&lds[cur][1][ABLK + wn * 4 + j][16 * lane];
        }
        if (kt * KC + KC > K) {
            // last chunk: k past K reads whatever follows in the plane row; zero it on both sides
            const int kv = K - kt * KC - 8 * lkg;
#pragma unroll
            for (int e = 0; e < 8; ++e)
                if (e >= kv) {
#pragma unroll
                    for (int i = 0; i < WMB; ++i) ah[i][e] = al[i][e] = (__bf16)0.0f;
#pragma unroll
                    for (int j = 0; j < 4; ++j) bh[j][e] = bl[j][e] = (__bf16)0.0f;
                }
        }
#pragma unroll
        for (int i = 0; i < WMB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        if ((kt % kFlush) == kFlush - 1) {
#pragma unroll
            for (int i = 0; i < WMB; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    sum[i][j] += acc[i][j];
                    acc[i][j] = f32x4{0, 0, 0, 0};
                }
        }
    }

    const float alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < WMB; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const f32x4 tot = sum[i][j] + acc[i][j];
            const int col = n0 + wn * 64 + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = m0 + wm * 16 * WMB + i * 16 + 4 * (lane >> 4) + t;
                if (row >= M) continue;
                if (tri && row + g.tri_off < col) continue;
                float* c = s.C + row + (long)col * s.ldc;
                float v = alpha * tot[t];
                if (beta != 0.0f) v += beta * *c;
                *c = v;
            }
        }
    }
}

// Planes of a column-major FP32 region: T = 1 -> plane row r, k = c; T = 0 -> plane row c, k = r.
// A 64 x 64 tile passes through LDS so both the FP32 reads and the plane writes are coalesced.
template <int T>
__global__ void split_planes_kernel(const float* __restrict__ A, long lda, int rows, int cols,
                                    unsigned short* __restrict__ hi, long ld, long ps) {
    __shared__ float tile[64][65];
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    if (T == 0) {
        // same orientation: element (r, c) -> hi[c * ld + r]; read and write down columns
        for (int c = ty; c < 64; c += 4) {
            const int gr = r0 + tx, gc = c0 + c;
            if (gr < rows && gc < cols) {
                const float v = A[gr + (long)gc * lda];
                const __bf16 h = (__bf16)v;
                const __bf16 l = (__bf16)(v - (float)h);
                hi[(long)gc * ld + gr] = __builtin_bit_cast(unsigned short, h);
                hi[ps + (long)gc * ld + gr] = __builtin_bit_cast(unsigned short, l);
            }
        }
        return;
    }
    for (int c = ty; c < 64; c += 4) {
        const int gr = r0 + tx, gc = c0 + c;
        tile[c][tx] = (gr < rows && gc < cols) ? A[gr + (long)gc * lda] : 0.0f;
    }
    __syncthreads();
    // transposed: element (r, c) -> hi[r * ld + c]; lanes run along c
    for (int r = ty; r < 64; r += 4) {
        const int gr = r0 + r, gc = c0 + tx;
        if (gr < rows && gc < cols) {
            const float v = tile[tx][r];
            const __bf16 h = (__bf16)v;
            const __bf16 l = (__bf16)(v - (float)h);
            hi[(long)gr * ld + gc] = __builtin_bit_cast(unsigned short, h);
            hi[ps + (long)gr * ld + gc] = __builtin_bit_cast(unsigned short, l);
        }
    }
}

// out[b][r + 128c] = sym(B)[128b + r][128b + c] from the lower triangle of B.
template <typename T>
__global__ void symdiag_kernel(const T* __restrict__ B, long ldb, int K, T* __restrict__ out) {
    const int idx = blockIdx.x * 256 + threadIdx.x;  // element of the 128 x 128 block
    const int r = idx % BN, c = idx / BN;
    const long gr = (long)blockIdx.y * BN + r, gc = (long)blockIdx.y * BN + c;
    if (gr >= K || gc >= K) return;
    out[(long)blockIdx.y * BN * BN + idx] = r >= c ? B[gr + gc * ldb] : B[gc + gr * ldb];
}

template <typename T>
int launch(const GemmArgs<T>& g, hipStream_t stream) {
    if (g.nseg == 0) return MOAO_OK;
    if (g.nseg < 0 || g.nseg > kMaxSeg || g.N < 0 || g.K < 0 || g.transb < 0 || g.transb > 2) {
        set_error("gemm: invalid arguments");
        return MOAO_ERR_ARG;
    }
    int maxM = 0;
    for (int i = 0; i < g.nseg; ++i) {
        maxM = g.seg[i].M > maxM ? g.seg[i].M : maxM;
        if (g.seg[i].B && g.transb == 2) {
            set_error("gemm: a symmetric B is shared by all segments");
            return MOAO_ERR_ARG;
        }
    }
    if (maxM == 0 || g.N == 0) return MOAO_OK;
    dim3 grid((maxM + BM - 1) / BM, (g.N + BN - 1) / BN, g.nseg);
    if (grid.y > 65535) {
        set_error("gemm: N too large for one launch");
        return MOAO_ERR_ARG;
    }
    double work = 0;
    if (profiling()) {
        for (int i = 0; i < g.nseg; ++i) {
            double area = (double)g.seg[i].M * g.N;
            if ((i == 0 || g.tri_all) && g.tri) {  // elements with row + tri_off >= col
                area = 0;
                for (int c = 0; c < g.N; ++c)
                    area += std::max(0, g.seg[i].M - std::max(0, c - g.tri_off));
            }
            work += 2.0 * area * g.K;
        }
    }
    ProfScope ps(sizeof(T) == 8 ? kProfGemmF64 : kProfGemmF32, work, stream);
    // 128-row tiles when they fill the device (>= 2 per CU), else 64-row tiles
    const bool small = (long)grid.x * grid.y * grid.z < 2L * 256;
    if constexpr (std::is_same<T, float>::value) {
        if (g.split == 2) {
            if (g.transb != 1 && g.transb != 0) {
                set_error("gemm: the planes product takes op(B) = B or B^T only");
                return MOAO_ERR_ARG;
            }
            for (int i = 0; i < g.nseg; ++i)
                if (g.seg[i].B) {
                    set_error("gemm: the planes product takes one B for all segments");
                    return MOAO_ERR_ARG;
                }
            if (small) {
                grid.x = (maxM + 63) / 64;
                hipLaunchKernelGGL((gemm_split2_kernel<64>), grid, dim3(NT), 0, stream, g);
            } else {
                hipLaunchKernelGGL((gemm_split2_kernel<128>), grid, dim3(NT), 0, stream, g);
            }
            MOAO_HIP_TRY(hipGetLastError());
            return MOAO_OK;
        }
        if (g.split) {
            set_error("gemm: unknown split mode");
            return MOAO_ERR_ARG;
        }
    }
    // the K x N slab loads address 128 columns with 32-bit byte offsets
    const bool kn_offsets_fit = (long)BN * g.ldb * (long)sizeof(T) < (1L << 31);
    if (small) {
        grid.x = (maxM + 63) / 64;
        if (g.transb == 0 && kn_offsets_fit)
            hipLaunchKernelGGL((gemm_dl_kernel<T, 0, kStages, 64>), grid, dim3(NT), 0, stream, g);
        else if (g.transb == 2 && g.symdiag && kn_offsets_fit)
            hipLaunchKernelGGL((gemm_dl_kernel<T, 2, kStages, 64>), grid, dim3(NT), 0, stream, g);
        else if (g.transb == 0)
            hipLaunchKernelGGL((gemm_kernel<T, 0, 2>), grid, dim3(NT), 0, stream, g);
        else if (g.transb == 1)
            hipLaunchKernelGGL((gemm_dl_kernel<T, 1, kStages, 64>), grid, dim3(NT), 0, stream, g);
        else
            hipLaunchKernelGGL((gemm_kernel<T, 2, 2>), grid, dim3(NT), 0, stream, g);
    } else if (g.transb == 0 && kn_offsets_fit)
        hipLaunchKernelGGL((gemm_dl_kernel<T, 0, kStages, 128>), grid, dim3(NT), 0, stream, g);
    else if (g.transb == 2 && g.symdiag && kn_offsets_fit)
        hipLaunchKernelGGL((gemm_dl_kernel<T, 2, kStages, 128>), grid, dim3(NT), 0, stream, g);
    else if (g.transb == 0)
        hipLaunchKernelGGL((gemm_kernel<T, 0, 4>), grid, dim3(NT), 0, stream, g);
    else if (g.transb == 1)
        hipLaunchKernelGGL((gemm_dl_kernel<T, 1, kStages, 128>), grid, dim3(NT), 0, stream, g);
    else
        hipLaunchKernelGGL((gemm_kernel<T, 2, 4>), grid, dim3(NT), 0, stream, g);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

}  // namespace

int gemm_f64(const GemmArgs<double>& a, hipStream_t stream) { return launch(a, stream); }
int gemm_f32(const GemmArgs<float>& a, hipStream_t stream) { return launch(a, stream); }

int split_planes(const float* A, long lda, int rows, int cols, unsigned short* hi, long ld,
                 long ps, int transposed, hipStream_t stream) {
    if (rows <= 0 || cols <= 0) return MOAO_OK;
    ProfScope pscope(kProfConvert, 8.0 * rows * cols, stream);
    const dim3 grid((rows + 63) / 64, (cols + 63) / 64);
    if (transposed)
        hipLaunchKernelGGL(split_planes_kernel<1>, grid, dim3(256), 0, stream, A, lda, rows, cols,
                           hi, ld, ps);
    else
        hipLaunchKernelGGL(split_planes_kernel<0>, grid, dim3(256), 0, stream, A, lda, rows, cols,
                           hi, ld, ps);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

size_t gemm_symdiag_elems(long K) { return (size_t)((K + BN - 1) / BN) * BN * BN; }

int gemm_symdiag(const double* B, long ldb, int K, double* out, hipStream_t stream) {
    if (K <= 0) return MOAO_OK;
    ProfScope ps(kProfOther, 0.0, stream);
    hipLaunchKernelGGL(symdiag_kernel<double>, dim3(BN * BN / 256, (K + BN - 1) / BN), dim3(256),
                       0, stream, B, ldb, K, out);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

}  // namespace moao

extern "C" {

// Plain device GEMM entry (utility; the stages above use the engine directly):
// C = alpha*A*op(B) + beta*C, op(B) = B^T (transb=1, B is N x K) or B (transb=0, B is K x N).
int moao_hip_dgemm_dev(int transb, int M, int N, int K, double alpha, const double* A, long lda,
                       const double* B, long ldb, double beta, double* C, long ldc, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (M < 0 || N < 0 || K < 0 || lda < std::max(1, M) || ldc < std::max(1, M) ||
        ldb < std::max(1, transb == 1 ? N : K) || transb < 0 || transb > 2 ||
        (transb == 2 && K != N)) {
        moao::set_error("dgemm: invalid arguments");
        return MOAO_ERR_ARG;
    }
    auto a = moao::gemm_args<double>(M, N, K, alpha, A, lda, B, ldb, transb, beta, C, ldc);
    hipStream_t s = moao::pick_stream(stream);
    if (transb != 2 || K == 0) return moao::gemm_f64(a, s);
    // symmetric B: its diagonal blocks in stream-ordered scratch (register-staged kernel if
    // that allocation fails)
    double* sd = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&sd), sizeof(double) * moao::gemm_symdiag_elems(K),
                       s) != hipSuccess) {
        (void)hipGetLastError();
        return moao::gemm_f64(a, s);
    }
    int rc = moao::gemm_symdiag(B, ldb, K, sd, s);
    if (rc == MOAO_OK) {
        a.symdiag = sd;
        rc = moao::gemm_f64(a, s);
    }
    (void)hipFreeAsync(sd, s);
    return rc;
}

int moao_hip_sgemm_dev(int transb, int M, int N, int K, float alpha, const float* A, long lda,
                       const float* B, long ldb, float beta, float* C, long ldc, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (M < 0 || N < 0 || K < 0 || lda < std::max(1, M) || ldc < std::max(1, M) ||
        ldb < std::max(1, transb == 1 ? N : K) || transb < 0 || transb > 2 ||
        (transb == 2 && K != N)) {
        moao::set_error("sgemm: invalid arguments");
        return MOAO_ERR_ARG;
    }
    auto a = moao::gemm_args<float>(M, N, K, alpha, A, lda, B, ldb, transb, beta, C, ldc);
    return moao::gemm_f32(a, moao::pick_stream(stream));
}

// The same product on the split-BF16 kernel (op(B) = B or B^T): what the mixed BF16
// reconstructor's updates run.
int moao_hip_sgemm_bf16x3_dev(int transb, int M, int N, int K, float alpha, const float* A,
                              long lda, const float* B, long ldb, float beta, float* C, long ldc,
                              void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (M < 0 || N < 0 || K < 0 || lda < std::max(1, M) || ldc < std::max(1, M) ||
        ldb < std::max(1, transb == 1 ? N : K) || transb < 0 || transb > 1) {
        moao::set_error("sgemm_bf16x3: invalid arguments");
        return MOAO_ERR_ARG;
    }
    auto a = moao::gemm_args<float>(M, N, K, alpha, A, lda, B, ldb, transb, beta, C, ldc);
    hipStream_t s = moao::pick_stream(stream);
    if (M == 0 || N == 0) return MOAO_OK;
    // both operands split into planes (stream-ordered scratch), then the planes kernel
    const long ld = ((long)std::max(K, 1) + 63) / 64 * 64;
    const long psa = (long)M * ld, psb = (long)N * ld;
    unsigned short* pl = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&pl), sizeof(unsigned short) * 2 * (psa + psb),
                       s) != hipSuccess) {
        (void)hipGetLastError();
        moao::set_error("sgemm_bf16x3: plane allocation failed");
        return MOAO_ERR_NOMEM;
    }
    unsigned short* pa = pl;
    unsigned short* pb = pl + 2 * psa;
    int rc = moao::split_planes(A, lda, M, K, pa, ld, psa, 1, s);
    if (rc == MOAO_OK)
        rc = transb == 1 ? moao::split_planes(B, ldb, N, K, pb, ld, psb, 1, s)
                         : moao::split_planes(B, ldb, K, N, pb, ld, psb, 0, s);
    if (rc == MOAO_OK) {
        a.split = 2;
        a.seg[0].ap = moao::Planes{pa, ld, psa};
        a.bp = moao::Planes{pb, ld, psb};
        rc = moao::gemm_f32(a, s);
    }
    (void)hipFreeAsync(pl, s);
    return rc;
}

}  // extern "C"
