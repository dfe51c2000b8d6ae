// x-derivative as a dense precomputed operator applied with MFMA.
//
// The compact first derivative of a periodic line with the block closure of
// one rank is d = A^-1 (B f): A the tridiagonal of initialize.c:633-654, B the
// 7-point right-hand side of mathlib.c:46-86 whose halo entries wrap
// (parallel.c:13-24).  M = A^-1 B is formed once on the host in FP64 by
// running the Thomas solve (the factors of GridTables) on every column of B,
// and is stored split into BF16 planes, M ~ Mh + Mm (+ Ml), each plane the
// BF16 rounding of what the previous planes leave.  The derivative of field
// v is then the GEMM D_v = M U_v over all y-lines at once (U_v: nx x ny,
// row-major, the field's own layout), computed with
// v_mfma_f32_16x16x16bf16 and FP32 accumulation:
//   NS = 3:  Mh Uh + Mh Um + Mm Uh                         (~2^-16 per term)
//   NS = 6:  + Mh Ul + Ml Uh + Mm Um                       (~2^-24 per term)
// with U split the same way on the fly.
//
// M decays by ~0.45 per row away from the diagonal (the filter-free compact
// scheme), so beyond ~120 rows its entries are below the BF16 range and all
// planes hold exact zeros there.  Each 64-row tile keeps the list of 16-column
// panels in which any plane is non-zero, and only those are multiplied: the
// product is the full dense one of the split operator.
#include "mhd2d_internal.h"
#include "mhd2d_tables.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace mhd2d {

namespace {

// round-to-nearest-even FP32 -> BF16 bits
inline unsigned short bf16_bits_host(float x)
{
    unsigned u;
    std::memcpy(&u, &x, 4);
    const unsigned r = ((u >> 16) & 1u) + 0x7fffu;
    return (unsigned short)((u + r) >> 16);
}

inline double bf16_value_host(unsigned short b)
{
    unsigned u = (unsigned)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

hipError_t XGemmOp::build(const GridTables& tb, int nsplit_)
{
    release();
    nx = tb.nx;
    ny = tb.ny;
    nsplit = nsplit_;
    if ((nsplit != 3 && nsplit != 6) || nx % 64 || ny % 64 || tb.xblocks != 1)
        return hipErrorInvalidValue;
    const int N = nx;
    const XStencil& st = tb.st[XD1];
    // host copies of the Thomas factors of the first derivative
    std::vector<double> l(N), invd(N), q(N);
    const SchemeAlphas al;
    GridTables::factor(N, al.alpha1, al.alphaS1, l.data(), invd.data(), q.data());
    // B: row i of the right-hand side over the wrapped line
    auto addB = [&](std::vector<double>& col, int k) {
        // column k of B: every row i whose stencil touches f[k]
        std::fill(col.begin(), col.end(), 0.0);
        auto wrap = [&](int i) { return ((i % N) + N) % N; };
        for (int i = 1; i <= N - 2; i++) {
            for (int m = 1; m <= 3; m++) {
                const double c = m == 1 ? st.c1 : (m == 2 ? st.c2 : st.c3);
                if (wrap(i + m) == k) col[i] += c;
                if (wrap(i - m) == k) col[i] -= c;
            }
        }
        for (int e = 0; e < 5; e++) {
            if (wrap(e - 2) == k) col[0] += st.e[e];                    // row 0 over f[-2..2]
            if (wrap(N + 1 - e) == k) col[N - 1] += st.esign * st.e[e];  // row N-1 over f[N+1..N-3]
        }
    };
    std::vector<double> M((size_t)N * N), col(N), y(N);
    for (int k = 0; k < N; k++) {
        addB(col, k);
        // Thomas with the factors (same recurrences as the device solve)
        double yy = 0.0;
        for (int i = 0; i < N; i++) {
            yy = col[i] - l[i] * yy;
            y[i] = yy;
        }
        double xx = 0.0;
        for (int i = N - 1; i >= 0; i--) {
            xx = y[i] * invd[i] + q[i] * xx;
            M[(size_t)i * N + k] = xx;
        }
    }
    // split into BF16 planes
    const size_t NN = (size_t)N * N;
    std::vector<unsigned short> planes((size_t)3 * NN, 0);
    for (size_t p = 0; p < NN; p++) {
        double r = M[p];
        for (int s = 0; s < 3; s++) {
            const unsigned short b = bf16_bits_host((float)r);
            planes[s * NN + p] = b;
            r -= bf16_value_host(b);
        }
    }
    // non-zero 16-column panels of every 64-row tile
    const int ntile = N / 64, npanel = N / 16;
    std::vector<std::vector<int>> lists(ntile);
    const int used = nsplit == 3 ? 2 : 3;
    kmax = 0;
    for (int t = 0; t < ntile; t++) {
        for (int pnl = 0; pnl < npanel; pnl++) {
            bool nz = false;
            for (int i = t * 64; i < t * 64 + 64 && !nz; i++)
                for (int k = pnl * 16; k < pnl * 16 + 16 && !nz; k++)
                    for (int s = 0; s < used && !nz; s++)
                        nz = (planes[s * NN + (size_t)i * N + k] & 0x7fff) != 0;
            if (nz) lists[t].push_back(pnl);
        }
        kmax = std::max(kmax, (int)lists[t].size());
    }
    panels_total = 0;
    std::vector<int> klist((size_t)ntile * kmax, 0), kcount(ntile);
    for (int t = 0; t < ntile; t++) {
        kcount[t] = (int)lists[t].size();
        panels_total += kcount[t];
        for (int q2 = 0; q2 < kcount[t]; q2++) klist[(size_t)t * kmax + q2] = lists[t][q2];
    }
    const size_t pbytes = (size_t)used * NN * sizeof(unsigned short);
    const size_t lbytes = (klist.size() + kcount.size()) * sizeof(int);
    hipError_t e = hipMalloc(&blob_, pbytes + lbytes);
    if (e != hipSuccess) return e;
    e = hipMemcpy(blob_, planes.data(), pbytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    char* base = (char*)blob_;
    e = hipMemcpy(base + pbytes, klist.data(), klist.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    e = hipMemcpy(base + pbytes + klist.size() * sizeof(int), kcount.data(),
                  kcount.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const unsigned short* pl = (const unsigned short*)blob_;
    mh = pl;
    mm = pl + NN;
    ml = used == 3 ? pl + 2 * NN : nullptr;
    klist_d = (const int*)(base + pbytes);
    kcount_d = klist_d + klist.size();
    bytes = pbytes + lbytes;
    return hipSuccess;
}

void XGemmOp::release()
{
    if (blob_) (void)hipFree(blob_);
    blob_ = nullptr;
    mh = mm = ml = nullptr;
    klist_d = kcount_d = nullptr;
}

XGemmOp::~XGemmOp() { release(); }

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_bits(float x)
{
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + (((u >> 16) & 1u) + 0x7fffu)) >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short b)
{
    return __uint_as_float((unsigned)b << 16);
}

// One workgroup: a 64 x 64 tile of D_v for one field of one instance; four
// waves of 32 x 32 (2 x 2 blocks of 16 x 16).  MFMA 16x16x16 operand layout:
// A[i][k] in lane (i + 16 * (k / 4)), element k % 4; B[k][j] in lane
// (j + 16 * (k / 4)), element k % 4; D[i][j] in lane (j + 16 * (i / 4)),
// element i % 4.
template <typename TI, int NS>
__global__ void __launch_bounds__(256) k_xgemm(XGemmArgs<TI> a)
{
    const int z = blockIdx.z;
    if (a.rs && a.rs[z].stop) return;
    const int v = blockIdx.y;
    const int ny = a.ny;
    const int jtiles = ny / 64;
    const int it = blockIdx.x / jtiles, jt = blockIdx.x % jtiles;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wi = w >> 1, wj = w & 1;
    const int row0 = it * 64 + wi * 32, col0 = jt * 64 + wj * 32;
    const int lr = lane & 15, lk = (lane >> 4) * 4;
    const TI* __restrict__ U = a.in[v] + z * a.bs_in;
    float* __restrict__ D = a.out[v] + z * a.bs_out;
    const size_t N = (size_t)a.nx;
    // column constant (the operator annihilates constants)
    TI c0[2];
#pragma unroll
    for (int cb = 0; cb < 2; cb++) c0[cb] = U[col0 + cb * 16 + lr];

    f4v acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++) acc[rb][cb] = f4v{0.f, 0.f, 0.f, 0.f};

    const int np = a.kcount[it];
    const int* __restrict__ kl = a.klist + (size_t)it * a.kmax;
    for (int qn = 0; qn < np; qn++) {
        const int k0 = kl[qn] * 16 + lk;
        s4v ah[2], am[2], al[2];
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            const size_t off = (size_t)(row0 + rb * 16 + lr) * N + k0;
            ah[rb] = *reinterpret_cast<const s4v*>(a.mh + off);
            am[rb] = *reinterpret_cast<const s4v*>(a.mm + off);
            if constexpr (NS == 6) al[rb] = *reinterpret_cast<const s4v*>(a.ml + off);
        }
        s4v bh[2], bm[2], bl[2];
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float x = (float)(U[(size_t)(k0 + e) * ny + col0 + cb * 16 + lr] - c0[cb]);
                const unsigned short h = bf16_bits(x);
                const float r1 = x - bf16_float(h);
                const unsigned short m = bf16_bits(r1);
                bh[cb][e] = (short)h;
                bm[cb][e] = (short)m;
                if constexpr (NS == 6) bl[cb][e] = (short)bf16_bits(r1 - bf16_float(m));
            }
        }
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                // small terms first
                if constexpr (NS == 6) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am[rb], bm[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[rb], bh[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[rb], bl[cb], acc[rb][cb], 0, 0, 0);
                }
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am[rb], bh[cb], acc[rb][cb], 0, 0, 0);
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[rb], bm[cb], acc[rb][cb], 0, 0, 0);
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[rb], bh[cb], acc[rb][cb], 0, 0, 0);
            }
    }
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                D[(size_t)(row0 + rb * 16 + lk + e) * ny + col0 + cb * 16 + lr] = acc[rb][cb][e];
}

template <typename TI>
hipError_t launch_xgemm(const XGemmArgs<TI>& a, int nv, hipStream_t s, int nb)
{
    if (a.nx % 64 || a.ny % 64 || nv < 1 || nv > 16 || nb < 1 || nb > 65535 || !a.mh || !a.mm)
        return hipErrorInvalidValue;
    dim3 grid((a.nx / 64) * (a.ny / 64), nv, nb), block(256);
    if (a.nsplit == 6) {
        if (!a.ml) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_xgemm<TI, 6>), grid, block, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_xgemm<TI, 3>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

template hipError_t launch_xgemm<double>(const XGemmArgs<double>&, int, hipStream_t, int);
template hipError_t launch_xgemm<float>(const XGemmArgs<float>&, int, hipStream_t, int);

}  // namespace mhd2d
