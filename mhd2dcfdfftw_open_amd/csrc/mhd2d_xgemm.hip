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
// row-major, the field's own layout), computed with gfx950's
// v_mfma_f32_16x16x32_bf16 and FP32 accumulation:
//   NS = 3:  Mh Uh + Mh Um + Mm Uh                         (~2^-16 per term)
//   NS = 6:  + Mh Ul + Ml Uh + Mm Um                       (~2^-24 per term)
// with U split the same way on the fly, once per element and workgroup.
//
// M decays by ~0.45 per row away from the diagonal (the filter-free compact
// scheme), so beyond ~120 rows its entries are below the BF16 range and all
// planes hold exact zeros there.  Each 64-row tile keeps the list of 32-column
// chunks in which any plane is non-zero, and only those are multiplied: the
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
    // non-zero 32-column chunks of every 64-row tile
    const int ntile = N / 64, npanel = N / XGEMM_KC;
    std::vector<std::vector<int>> lists(ntile);
    const int used = nsplit == 3 ? 2 : 3;
    kmax = 0;
    for (int t = 0; t < ntile; t++) {
        for (int pnl = 0; pnl < npanel; pnl++) {
            bool nz = false;
            for (int i = t * 64; i < t * 64 + 64 && !nz; i++)
                for (int k = pnl * XGEMM_KC; k < pnl * XGEMM_KC + XGEMM_KC && !nz; k++)
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
    // 128- and 256-row tiles (2 or 4 consecutive 64-row tiles): the union of
    // their chunk lists, each entry (chunk << 4) | (bit h: the 64-row
    // sub-tile h multiplies it)
    auto group_lists = [&](int R, std::vector<int>& flat, std::vector<int>& count) {
        const int nt = ntile / R;
        std::vector<std::vector<int>> lr(nt);
        int km = 0;
        for (int t = 0; t < nt; t++) {
            std::vector<int> mask(npanel, 0);
            for (int h = 0; h < R; h++)
                for (int c : lists[R * t + h]) mask[c] |= 1 << h;
            for (int c = 0; c < npanel; c++)
                if (mask[c]) lr[t].push_back((c << 4) | mask[c]);
            km = std::max(km, (int)lr[t].size());
        }
        flat.assign((size_t)std::max(1, nt * km), 0);
        count.assign(std::max(1, nt), 0);
        for (int t = 0; t < nt; t++) {
            count[t] = (int)lr[t].size();
            for (int q2 = 0; q2 < count[t]; q2++) flat[(size_t)t * km + q2] = lr[t][q2];
        }
        return km;
    };
    std::vector<int> klist2, kcount2, klist4, kcount4;
    kmax2 = group_lists(2, klist2, kcount2);
    kmax4 = group_lists(4, klist4, kcount4);
    // one int array: klist, klist2, klist4, kcount, kcount2, kcount4
    const size_t o2 = klist.size(), o4 = o2 + klist2.size(), oc = o4 + klist4.size();
    const size_t oc2 = oc + kcount.size(), oc4 = oc2 + kcount2.size();
    std::vector<int> ints;
    for (const auto* v : {&klist, &klist2, &klist4, &kcount, &kcount2, &kcount4})
        ints.insert(ints.end(), v->begin(), v->end());
    const size_t pbytes = (size_t)used * NN * sizeof(unsigned short);
    const size_t lbytes = ints.size() * sizeof(int);
    hipError_t e = hipMalloc(&blob_, pbytes + lbytes);
    if (e != hipSuccess) return e;
    e = hipMemcpy(blob_, planes.data(), pbytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    char* base = (char*)blob_;
    e = hipMemcpy(base + pbytes, ints.data(), lbytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const unsigned short* pl = (const unsigned short*)blob_;
    mh = pl;
    mm = pl + NN;
    ml = used == 3 ? pl + 2 * NN : nullptr;
    const int* id = (const int*)(base + pbytes);
    klist_d = id;
    klist2_d = ntile >= 2 ? id + o2 : nullptr;
    klist4_d = ntile >= 4 ? id + o4 : nullptr;
    kcount_d = id + oc;
    kcount2_d = ntile >= 2 ? id + oc2 : nullptr;
    kcount4_d = ntile >= 4 ? id + oc4 : nullptr;
    bytes = pbytes + lbytes;
    return hipSuccess;
}

void XGemmOp::release()
{
    if (blob_) (void)hipFree(blob_);
    blob_ = nullptr;
    mh = mm = ml = nullptr;
    klist_d = kcount_d = klist2_d = kcount2_d = klist4_d = kcount4_d = nullptr;
}

XGemmOp::~XGemmOp() { release(); }

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

// Two FP32 values split into BF16 planes, round to nearest even at each
// level: p[0] their BF16 roundings, p[1] the roundings of what p[0] leaves,
// p[2] (NP == 3) of what p[0] + p[1] leave; each plane packed in one dword
// (first value in the low half).  The casts compile to gfx950's
// v_cvt_pk_bf16_f32 (one instruction per plane and pair).
template <int NP>
__device__ __forceinline__ void bf16_split2(float x0, float x1, unsigned (&p)[3])
{
    const bf2v h = {(__bf16)x0, (__bf16)x1};
    const float r0 = x0 - (float)h[0], r1 = x1 - (float)h[1];
    const bf2v m = {(__bf16)r0, (__bf16)r1};
    p[0] = __builtin_bit_cast(unsigned, h);
    p[1] = __builtin_bit_cast(unsigned, m);
    if constexpr (NP == 3) {
        const bf2v l = {(__bf16)(r0 - (float)m[0]), (__bf16)(r1 - (float)m[1])};
        p[2] = __builtin_bit_cast(unsigned, l);
    }
}

__device__ __forceinline__ f4v mfma_bf16(const s8v& a, const s8v& b, const f4v& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a),
                                                   __builtin_bit_cast(bf8v, b), c, 0, 0, 0);
}

// Workgroup: a (64 RH)-row x 256-column tile of D_v (one field v, one
// instance).  RH = 1: 4 waves of 64 x 64; RH = 2: 8 waves of 64 x 64; RH = 4:
// 8 waves of 128 x 64 (4 x 4 or 8 x 4 blocks of 16 x 16 accumulators).  For
// every 32-column chunk of M's row tile that is non-zero in any of its
// 64-row sub-tiles:
//  * the 32 x 256 block of U (rows = the chunk's columns) is read once by
//    the workgroup (each column by RH threads), the column's row-0 value
//    subtracted (the operator annihilates constants), split into BF16
//    planes and written to LDS transposed, [plane][column][k]; the
//    (64 RH) x 32 block of each M plane goes to LDS as [plane][row][k].
//    Rows and columns are padded to 40 elements (80 bytes: the 16-byte
//    fragment reads of 16 consecutive rows or columns then cover all 64
//    banks once);
//  * each wave reads its A and B fragments from LDS and runs the split
//    products of its sub-tiles that have the chunk, small terms first;
//  * the next chunk's U and M blocks are loaded into registers while the
//    MFMAs run.
// Taller tiles share a U block among more rows: against 64-row tiles the
// staging per output element (loads, split, LDS writes) falls by about a
// third with 128 rows at nx = 4096, by two thirds with 256.
// MFMA 16x16x32 operands: lane (r = lane % 16, g = lane / 16) holds 8
// consecutive k of row r (A) or column r (B), k = 8 g .. 8 g + 7; the same k
// mapping on both sides pairs the products correctly.  D[i][j] is in lane
// (j + 16 * (i / 4)), element i % 4.
template <typename TI, int NS, int RH>
__global__ void __launch_bounds__(RH == 1 ? 256 : 512, RH == 1 ? 2 : 1) k_xgemm(XGemmArgs<TI> a)
{
    constexpr int NT = RH == 1 ? 256 : 512;  // threads: 4 or 8 waves
    constexpr int WRB = RH == 4 ? 8 : 4;     // 16-row blocks of a wave
    constexpr int WSUB = WRB / 4;            // 64-row sub-tiles of a wave
    constexpr int NP = NS == 6 ? 3 : 2;  // planes of U and of M
    constexpr int CT = 256;              // columns per workgroup
    constexpr int KC = XGEMM_KC;         // k per chunk
    constexpr int KT = KC * 256 / NT;    // k of a column staged by one thread
    constexpr int CS = 40;               // LDS row / column stride (bf16 elements)
    constexpr int RT = 64 * RH;          // rows per workgroup
    constexpr int MP = NP * RT * (KC / 8);  // 16-byte pieces of an M block
    constexpr int MPT = (MP + NT - 1) / NT; // per thread
    __shared__ __attribute__((aligned(16))) unsigned short su[NP * CT * CS];
    __shared__ __attribute__((aligned(16))) unsigned short sm[NP * RT * CS];
    const int z = blockIdx.z;
    if (a.rs && a.rs[z].stop) return;
    const int v = blockIdx.y;
    const int ny = a.ny;
    const int ctiles = (ny + CT - 1) / CT;
    const int it = blockIdx.x / ctiles, ct = blockIdx.x % ctiles;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int wh = w >> 2, wc = w & 3;        // this wave's row group and column block
    const int lr = lane & 15, lg = lane >> 4;
    const int row0 = it * RT;
    const int colw = ct * CT + wc * 64;       // this wave's first column
    const bool wact = colw < ny;              // waves beyond a narrow grid only stage
    const int mycol = ct * CT + (t % CT);     // the column this thread stages
    const int kpart = t / CT;                 // and its part of the chunk's k
    const bool cact = mycol < ny;
    const TI* __restrict__ U = a.in[v] + z * a.bs_in;
    float* __restrict__ D = a.out[v] + z * a.bs_out;
    const size_t N = (size_t)a.nx;
    const unsigned short* const mpl[3] = {a.mh, a.mm, a.ml};
    const TI c0 = cact ? U[mycol] : (TI)0;

    f4v acc[WRB][4];
#pragma unroll
    for (int rb = 0; rb < WRB; rb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++) acc[rb][cb] = f4v{0.f, 0.f, 0.f, 0.f};

    const int* __restrict__ kl = RH == 4   ? a.klist4 + (size_t)it * a.kmax4
                                 : RH == 2 ? a.klist2 + (size_t)it * a.kmax2
                                           : a.klist + (size_t)it * a.kmax;
    const int nch = RH == 4 ? a.kcount4[it] : (RH == 2 ? a.kcount2[it] : a.kcount[it]);
    // chunk index and the 64-row sub-tiles that multiply it
    auto chunk = [&](int q) { return RH == 1 ? kl[q] : (kl[q] >> 4); };
    auto mask = [&](int q) { return RH == 1 ? 1 : (kl[q] & 15); };
    TI pre[KT];
    s8v mpre[MPT];
    // this thread's U values and M pieces of chunk q into registers
    auto fetch = [&](int q) {
        const int k0 = chunk(q) * KC;
        // unconditional loads (a column beyond a narrow grid reads column 0
        // and is zeroed in the split): a conditional load per element became
        // a branch with a wait for each one
        const TI* col = U + (size_t)(k0 + kpart * KT) * ny + (cact ? mycol : 0);
#pragma unroll
        for (int e = 0; e < KT; e++) pre[e] = col[(size_t)e * ny];
#pragma unroll
        for (int i = 0; i < MPT; i++) {
            const int pc = t + NT * i;  // piece: plane, row, k group
            if (pc < MP) {
                const int pl = pc / (RT * (KC / 8)), r = (pc / (KC / 8)) % RT, kg = pc % (KC / 8);
                mpre[i] = *reinterpret_cast<const s8v*>(mpl[pl] + (size_t)(row0 + r) * N + k0 +
                                                        kg * 8);
            }
        }
    };
    if (nch > 0) fetch(0);
    for (int q = 0; q < nch; q++) {
        __syncthreads();  // the previous chunk's fragments have been read
        const int sc = t % CT;
#pragma unroll
        for (int g = 0; g < KT / 8; g++) {
            u4v h, m, l;
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                const float x0 = cact ? (float)(pre[g * 8 + e] - c0) : 0.f;
                const float x1 = cact ? (float)(pre[g * 8 + e + 1] - c0) : 0.f;
                unsigned p[3];
                bf16_split2<NP>(x0, x1, p);
                h[e / 2] = p[0];
                m[e / 2] = p[1];
                if constexpr (NS == 6) l[e / 2] = p[2];
            }
            const int ko = kpart * KT + g * 8;
            *reinterpret_cast<u4v*>(su + (0 * CT + sc) * CS + ko) = h;
            *reinterpret_cast<u4v*>(su + (1 * CT + sc) * CS + ko) = m;
            if constexpr (NS == 6) *reinterpret_cast<u4v*>(su + (2 * CT + sc) * CS + ko) = l;
        }
#pragma unroll
        for (int i = 0; i < MPT; i++) {
            const int pc = t + NT * i;
            if (pc < MP) {
                const int pl = pc / (RT * (KC / 8)), r = (pc / (KC / 8)) % RT, kg = pc % (KC / 8);
                *reinterpret_cast<s8v*>(sm + (pl * RT + r) * CS + kg * 8) = mpre[i];
            }
        }
        __syncthreads();
        // the wave's 64-row sub-tiles that multiply chunk q
        const int wmask = (mask(q) >> (wh * WSUB)) & ((1 << WSUB) - 1);
        const bool mine = wact && wmask != 0;
        if (q + 1 < nch) fetch(q + 1);  // in flight during the MFMAs
        if (mine) {
            s8v bh[4], bm[4], bl[4];
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                const int c = wc * 64 + cb * 16 + lr;
                bh[cb] = *reinterpret_cast<const s8v*>(su + (0 * CT + c) * CS + lg * 8);
                bm[cb] = *reinterpret_cast<const s8v*>(su + (1 * CT + c) * CS + lg * 8);
                if constexpr (NS == 6)
                    bl[cb] = *reinterpret_cast<const s8v*>(su + (2 * CT + c) * CS + lg * 8);
            }
#pragma unroll
            for (int rb = 0; rb < WRB; rb++) {
                if (WSUB > 1 && !((wmask >> (rb / 4)) & 1)) continue;
                const int r = wh * WRB * 16 + rb * 16 + lr;
                const s8v ah = *reinterpret_cast<const s8v*>(sm + (0 * RT + r) * CS + lg * 8);
                const s8v am = *reinterpret_cast<const s8v*>(sm + (1 * RT + r) * CS + lg * 8);
                s8v al;
                if constexpr (NS == 6)
                    al = *reinterpret_cast<const s8v*>(sm + (2 * RT + r) * CS + lg * 8);
#pragma unroll
                for (int cb = 0; cb < 4; cb++) {
                    // small terms first
                    if constexpr (NS == 6) {
                        acc[rb][cb] = mfma_bf16(am, bm[cb], acc[rb][cb]);
                        acc[rb][cb] = mfma_bf16(al, bh[cb], acc[rb][cb]);
                        acc[rb][cb] = mfma_bf16(ah, bl[cb], acc[rb][cb]);
                    }
                    acc[rb][cb] = mfma_bf16(am, bh[cb], acc[rb][cb]);
                    acc[rb][cb] = mfma_bf16(ah, bm[cb], acc[rb][cb]);
                    acc[rb][cb] = mfma_bf16(ah, bh[cb], acc[rb][cb]);
                }
            }
        }
    }
    if (!wact) return;
    // 16 lanes write 64 contiguous bytes of a row; nontemporal (the RHS
    // reads the derivatives once, later)
    const int rw = row0 + wh * WRB * 16;
#pragma unroll
    for (int rb = 0; rb < WRB; rb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                __builtin_nontemporal_store(
                    acc[rb][cb][e], D + (size_t)(rw + rb * 16 + lg * 4 + e) * ny + colw + cb * 16 + lr);
}

template <typename TI>
hipError_t launch_xgemm(const XGemmArgs<TI>& a, int nv, hipStream_t s, int nb)
{
    if (a.nx % 64 || a.ny % 64 || nv < 1 || nv > 16 || nb < 1 || nb > 65535 || !a.mh || !a.mm)
        return hipErrorInvalidValue;
    const int ctiles = (a.ny + 255) / 256;
    // Row tiles: 256 rows for the 3-product form (the U staging per output
    // halves against 128: 4096^2 mixed 0.71 -> 0.62 ms), 128 for 6 products
    // (256-row tiles need 256 VGPRs there and measured 1.3x slower in the
    // mixed form, 1.1x in FP32); lines that do not hold whole tiles take the
    // next smaller ones
    const int rh = (a.nsplit == 3 && a.nx % 256 == 0 && a.klist4 && a.kcount4) ? 4
                   : (a.nx % 128 == 0 && a.klist2 && a.kcount2) ? 2 : 1;
    dim3 grid((a.nx / (64 * rh)) * ctiles, nv, nb), block(rh == 1 ? 256 : 512);
    if (a.nsplit == 6 && !a.ml) return hipErrorInvalidValue;
    if (rh == 4) {
        hipLaunchKernelGGL((k_xgemm<TI, 3, 4>), grid, block, 0, s, a);
    } else if (rh == 2) {
        if (a.nsplit == 6) hipLaunchKernelGGL((k_xgemm<TI, 6, 2>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_xgemm<TI, 3, 2>), grid, block, 0, s, a);
    } else {
        if (a.nsplit == 6) hipLaunchKernelGGL((k_xgemm<TI, 6, 1>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_xgemm<TI, 3, 1>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

template hipError_t launch_xgemm<double>(const XGemmArgs<double>&, int, hipStream_t, int);
template hipError_t launch_xgemm<float>(const XGemmArgs<float>&, int, hipStream_t, int);

}  // namespace mhd2d
