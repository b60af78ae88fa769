// k_deform_adj.hip -- fused forward-energy + adjoint-gradient + Jacobi-update kernel.
//
// Discrete energy (no reference counterpart; its parts restate the reference's terms):
//   q_j   = j + u_j (index space), psi_j = T(q_j) with T the trilinear interpolant of the
//           live SDF under the reference's read convention (SDF.cpp:503-524: truncating base,
//           out-of-grid corners = Max + eps); g_j = grad T(q_j), H_j = its mixed second
//           derivatives (a trilinear has none on the diagonal).
//   m_j   = active mask, -clip <= psi_j <= Max - eps (KillingFusion.cpp:272-274).
//   E     = sum_j m_j [ (psi_j - phi_can_j)^2 / (2 V)            data  (KillingFusion.cpp:444-455)
//                     + omega_s (s_j - 1)^2 / 2 ]                 level set (KillingFusion.cpp:464-476)
//         + omega_k sum_{i in [-1,n-1]^3} ||D u_i||_F^2 + gamma tr((D u_i)^2)
//                                                                  Killing (DisplacementField.cpp:171-193)
//   s_j = sqrt(|g_j|^2 + eps^2), (D u_i)_ab = u^a_{i+e_b} - u^a_i, u = 0 outside the grid.
// dE/du_j = m_j [ (psi_j - phi_can_j)/V g_j + omega_s (s_j - 1)/s_j H_j g_j ]
//         - 2 omega_k [ Lap u_j + gamma sum_b (D_a u^b_j - D_a u^b_{j-e_b}) ]_a
//
// One pass per iteration reads u_j, phi_can_j, the 8 live-SDF corners around q_j and the
// 1-neighbourhood of u (12 neighbours), and writes u_j - alpha dE/du_j for active voxels.
// Compulsory HBM traffic per voxel-iteration: 12 + s_phi bytes (u and phi_live read), plus
// 12 + s_phi where a voxel can move (its u written, phi_can read).
//
// Three launch shapes share the per-voxel arithmetic (adj_voxel):
//  * k_adj_flat (Jacobi steps): a thread owns two groups of 4 consecutive voxels of the flat
//    index, 1024 voxels apart.  A step moves only active voxels: a voxel whose own u is zero
//    samples phi_live[j] exactly, so u, phi_live and u_out stream as 16-byte / 8-byte accesses
//    and only active voxels (and moved ones, to find out) run the stencil.  Grids of
//    3 N >= 2^30 take its form with 64-bit offsets.
//  * k_adj_rows (energy and gradient passes): a thread owns 4 consecutive x voxels of a row, a
//    wavefront 256 voxels of a row, a block 4 rows of one slice.  These passes need every voxel
//    whose stencil sees a nonzero u, and take that set from a bit mask over the 4-voxel groups
//    of each row.
//  * k_adj_any (energy and gradient passes on grids without a mask, 3 N >= 2^30): one voxel
//    per thread, neighbours read individually, 64-bit offsets.
// Blocks go to the XCDs in chunks of 64 consecutive tiles dealt round-robin: neighbouring
// tiles of a chunk share an XCD's L2, and the band's stencils spread over every XCD.
#include "kf_adjoint.hpp"

#include <type_traits>

namespace {

constexpr int VEC = 4;                  // voxels per thread along x in k_adj_rows
constexpr int ROW = 64 * VEC;           // voxels of a row segment per wavefront
constexpr int RPT = 1;                  // rows per thread in k_adj_rows (4 rows per block)
constexpr int ROWS = BLOCK / 64 * RPT;  // rows per block (a row tile)

// ---------------------------------------------------------------------------
// Any shape: one voxel per thread.
// ---------------------------------------------------------------------------

// The full stencil of voxel i = (x, y, z) with 64-bit offsets (any grid): as adj_job, which
// indexes with 32-bit offsets.  A step writes the moved u_i of an active voxel or, with
// WRITE_ALL, the new u_i of every voxel.
template <class S, int MODE, bool WRITE_ALL>
__device__ inline void adj_job_wide(const PhiRead<S> &L, const PhiRead<S> &Cn, const float *__restrict__ u_in,
                                    float *__restrict__ u_out, const AdjConsts &c, long long *div,
                                    float *__restrict__ e_voxel, int64_t i, int x, int y, int z, double &mx,
                                    double &esum, unsigned long long &act)
{
    const GridDims &g = L.g;
    Nbhd nb;
    nb.uc[0] = u_in[i];
    nb.uc[1] = u_in[g.n + i];
    nb.uc[2] = u_in[2 * g.n + i];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nb.up[0][a] = uval(u_in, g, a, x + 1, y, z);
        nb.um[0][a] = uval(u_in, g, a, x - 1, y, z);
        nb.up[1][a] = uval(u_in, g, a, x, y + 1, z);
        nb.um[1][a] = uval(u_in, g, a, x, y - 1, z);
        nb.up[2][a] = uval(u_in, g, a, x, y, z + 1);
        nb.um[2][a] = uval(u_in, g, a, x, y, z - 1);
    }
    nb.dgl[0][1] = uval(u_in, g, 1, x + 1, y - 1, z);
    nb.dgl[0][2] = uval(u_in, g, 2, x + 1, y, z - 1);
    nb.dgl[1][0] = uval(u_in, g, 0, x - 1, y + 1, z);
    nb.dgl[1][2] = uval(u_in, g, 2, x, y + 1, z - 1);
    nb.dgl[2][0] = uval(u_in, g, 0, x - 1, y, z + 1);
    nb.dgl[2][1] = uval(u_in, g, 1, x, y - 1, z + 1);
    int bx, by, bz;
    float fx, fy, fz, v[8];
    kf_split(x, nb.uc[0], bx, fx);
    kf_split(y, nb.uc[1], by, fy);
    kf_split(z, nb.uc[2], bz, fz);
    L.cell(bx, by, bz, v);
    VoxelOut o;
    adj_voxel<MODE == KF_MODE_ENERGY>(c, v, fx, fy, fz, c.use_data ? Cn.at(x, y, z) : 0.f, nb, x, y, z, o);
    if (o.active)
        act += 1;
    esum += (double)o.e;
    if (MODE == KF_MODE_GRAD) {
        u_out[i] = o.g[0];
        u_out[g.n + i] = o.g[1];
        u_out[2 * g.n + i] = o.g[2];
    } else if (MODE == KF_MODE_ENERGY) {
        if (e_voxel)
            e_voxel[i] = o.e;
    } else if (o.active || WRITE_ALL) {
        float n[3];
        adj_update(c, o, nb.uc, i, div, mx, n);
        u_out[i] = n[0];
        u_out[g.n + i] = n[1];
        u_out[2 * g.n + i] = n[2];
    }
}

template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void k_adj_any(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                   float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                   BlockPartial *partials, float *__restrict__ e_voxel, int parity,
                                                   const BlockPartial *prev, unsigned int prev_nb)
{
    if (!step_gate<MODE>(st, prev, prev_nb, parity))
        return;
    const GridDims g = L.g;
    const int64_t i = (int64_t)kf_xcd_chunk_tile<64>(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;
    if (i < g.n) {
        int x, y, z;
        kf_xyz(g, i, x, y, z);
        adj_job_wide<S, MODE, true>(L, Cn, u_in, u_out, c, &st->div_slot[parity], e_voxel, i, x, y, z, mx, esum,
                                    act);
    }
    if (MODE != KF_MODE_GRAD)
        kf_block_publish<BLOCK>(partials, blockIdx.x, esum, mx, act);
}

// ---------------------------------------------------------------------------
// Energy and gradient passes by rows: 4 voxels per thread, 16-byte row loads.
// ---------------------------------------------------------------------------

// Nonzero masks.  mask bit (row, group) = 1 iff some u component of the 4-voxel group
// (x = 4 group .. 4 group + 3) of that row may be nonzero.  A row (y, z) holds MW words per
// 256-voxel segment (one word per half wavefront), so each wavefront owns its two words and
// writes them with a ballot -- no atomics.
constexpr int MW = ROW / VEC / 32;

// Word wi of row (y, z) dilated by one group along x: bit b is set iff group 32 wi + b - 1,
// + b or + b + 1 of the row is; 0 outside the grid.
__device__ inline unsigned int row_word_dilated(const unsigned int *__restrict__ m, int wpr, const GridDims &g, int y,
                                                int z, int wi)
{
    if (y < 0 || y >= g.ny || z < 0 || z >= g.nz)
        return 0u;
    const unsigned int *row = m + ((unsigned int)z * (unsigned int)g.ny + (unsigned int)y) * (unsigned int)wpr;
    const unsigned int w = row[wi];
    const unsigned int lo = wi > 0 ? row[wi - 1] >> 31 : 0u;
    const unsigned int hi = wi + 1 < wpr ? row[wi + 1] << 31 : 0u;
    return w | (w << 1) | (w >> 1) | lo | hi;
}

// Stencil mask: bit (row, group) of d is set iff the nonzero mask m has a bit set among
// groups gx - 1 .. gx + 1 of row (y, z) and of the rows (y +- 1, z), (y, z +- 1),
// (y + 1, z - 1), (y - 1, z + 1) -- every u value the adjoint stencil of the group's four
// voxels reads.  One thread per word; the energy and gradient passes then test one bit.
__global__ __launch_bounds__(BLOCK) void k_adj_dilate(const unsigned int *__restrict__ m, GridDims g, int wpr,
                                                      unsigned int *__restrict__ d)
{
    const unsigned int words = (unsigned int)wpr * (unsigned int)g.ny * (unsigned int)g.nz;
    const unsigned int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= words)
        return;
    const int wi = (int)(k % (unsigned int)wpr);
    const unsigned int row = k / (unsigned int)wpr;
    const int y = (int)(row % (unsigned int)g.ny), z = (int)(row / (unsigned int)g.ny);
    d[k] = row_word_dilated(m, wpr, g, y, z, wi) | row_word_dilated(m, wpr, g, y + 1, z, wi) |
           row_word_dilated(m, wpr, g, y - 1, z, wi) | row_word_dilated(m, wpr, g, y, z + 1, wi) |
           row_word_dilated(m, wpr, g, y, z - 1, wi) | row_word_dilated(m, wpr, g, y + 1, z - 1, wi) |
           row_word_dilated(m, wpr, g, y - 1, z + 1, wi);
}

// Bit of group gx = 64 seg + lane of row (y, z) in stencil mask d (y, z in the grid).
__device__ inline unsigned int stencil_bit(const unsigned int *__restrict__ d, int wpr, const GridDims &g, int y,
                                           int z, int seg, int lane)
{
    const unsigned int *row = d + ((unsigned int)z * (unsigned int)g.ny + (unsigned int)y) * (unsigned int)wpr;
    const int gx = seg * (ROW / VEC) + lane;
    return (row[gx >> 5] >> (gx & 31)) & 1u;
}

// Row segments for rows that are not a multiple of 4 voxels.  A group then starts off a 16-byte
// boundary (read as one 4-byte-aligned 16-byte access), and the row's last group may end
// early: its elements past the row end read as zero, as row4's groups outside the grid do.
__device__ inline float4 row4u(const float *__restrict__ u, const GridDims &g, int c, int x0, int y, int z)
{
    if (!(x0 < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz))
        return make_float4(0.f, 0.f, 0.f, 0.f);
    const float *p = u + ((unsigned int)c * nvox32(g) + g.idx32(x0, y, z));
    if (x0 + VEC <= g.nx)
        return kf_load_stream4u(p);
    const int m = g.nx - x0;  // 1 .. 3 voxels left in the row
    return make_float4(p[0], m > 1 ? p[1] : 0.f, m > 2 ? p[2] : 0.f, 0.f);
}

template <bool UNAL>
__device__ inline float4 row_group(const float *__restrict__ u, const GridDims &g, int c, int x0, int y, int z)
{
    return UNAL ? row4u(u, g, c, x0, y, z) : row4(u, g, c, x0, y, z);
}

// Stores the group v at dst, whose row has left >= 1 voxels from dst on (the layout of
// row_group's reads).
template <bool UNAL>
__device__ inline void put_group(float *__restrict__ dst, const float4 &v, int left)
{
    if (!UNAL) {
        *reinterpret_cast<float4 *>(dst) = v;
    } else if (left >= VEC) {
        *reinterpret_cast<kf_f4u *>(dst) = kf_f4u{v.x, v.y, v.z, v.w};
    } else {
        dst[0] = v.x;
        if (left > 1)
            dst[1] = v.y;
        if (left > 2)
            dst[2] = v.z;
    }
}

// Writes the wavefront's mask words of row (y, z): bit lane = nz of that lane's group.
__device__ inline void put_bits(unsigned int *__restrict__ m, int wpr, const GridDims &g, int y, int z, int seg,
                                int lane, bool nz)
{
    const unsigned long long bal = __ballot(nz);
    if (lane == 0 && y < g.ny) {
        unsigned int *row = m + ((unsigned int)z * (unsigned int)g.ny + (unsigned int)y) * (unsigned int)wpr;
        row[seg * MW] = (unsigned int)bal;
        row[seg * MW + 1] = (unsigned int)(bal >> 32);
    }
}

// UNAL: the row length is not a multiple of 4, so a group starts off a 16-byte boundary (u
// groups move as 4-byte-aligned 16-byte accesses, phi_live element by element) and the last
// group of a row holds 1 .. 3 voxels.
template <class S, int MODE, bool UNAL>
__global__ __launch_bounds__(BLOCK) void k_adj_rows(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                    BlockPartial *partials, float *__restrict__ e_voxel, int nseg,
                                                    int nyt, const unsigned int *__restrict__ mask_in, int parity,
                                                    const BlockPartial *prev, unsigned int prev_nb)
{
    if (!step_gate<MODE>(st, prev, prev_nb, parity))
        return;
    const GridDims g = L.g;
    const unsigned int tile = kf_xcd_chunk_tile<64>(blockIdx.x, gridDim.x);
    const int seg = (int)(tile % (unsigned)nseg);
    const int yt = (int)((tile / (unsigned)nseg) % (unsigned)nyt);
    const int z = (int)(tile / ((unsigned)nseg * (unsigned)nyt));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x0 = seg * ROW + lane * VEC;
    const int wpr = nseg * MW;
    const unsigned int n32 = nvox32(g);
    int yr[RPT];
    bool valid[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        yr[r] = yt * ROWS + w + r * (BLOCK / 64);
        valid[r] = x0 < g.nx && yr[r] < g.ny;
    }

    // Phase 1 -- quiet test.  A voxel is quiet when every u value its stencil touches is
    // zero (everywhere away from the surface band); then psi_j = phi_live[j] exactly (zero
    // interpolation weights) and an inactive quiet voxel has zero gradient and energy.  The
    // stencil of a thread's 4 voxels reaches groups gx - 1 .. gx + 1 of seven rows; one bit of
    // the stencil mask (k_adj_dilate) holds their union, so quietness costs one load and no
    // u values beyond the thread's own.  All loads are issued up front.
    float4 C[RPT][3];
    float live4[RPT][VEC];
    unsigned int nb_bits[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int y = yr[r];
#pragma unroll
        for (int a = 0; a < 3; a++)
            C[r][a] = row_group<UNAL>(u_in, g, a, x0, y, z);
#pragma unroll
        for (int k = 0; k < VEC; k++)
            live4[r][k] = 0.f;
        if (valid[r] && !UNAL) {
            S::load4(L.p, (int64_t)g.idx32(x0, y, z), L.scale, live4[r]);
        } else if (valid[r] && x0 + VEC <= g.nx) {
            kf_load4_unaligned<S>(L.p + g.idx32(x0, y, z), L.scale, live4[r]);
        } else if (valid[r]) {
            const unsigned int i0 = g.idx32(x0, y, z);
#pragma unroll
            for (int k = 0; k < VEC - 1; k++)
                if (x0 + k < g.nx)
                    live4[r][k] = S::load(L.p, (int64_t)(i0 + (unsigned int)k), L.scale);
        }
        nb_bits[r] = 0u;
        if (MODE != KF_MODE_STEP && valid[r])
            nb_bits[r] = stencil_bit(mask_in, wpr, g, y, z, seg, lane);
    }
    // need bit r * VEC + k: voxel (x0 + k, yr[r]) goes to phase 2.  A Jacobi step moves only
    // active voxels, so it needs the voxels that are active or whose sample moved (u_j != 0,
    // activity unknown until the gather); energy and gradient passes also need every voxel
    // whose stencil sees a nonzero u.
    unsigned int need = 0u;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const bool quiet = nb_bits[r] == 0u && ((bits4(C[r][0]) | bits4(C[r][1]) | bits4(C[r][2])) & 0x7fffffffu) == 0u;
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            const bool moved = ((__float_as_uint(f4(C[r][0], k)) | __float_as_uint(f4(C[r][1], k)) |
                                 __float_as_uint(f4(C[r][2], k))) & 0x7fffffffu) != 0u;
            const bool in_band = !(live4[r][k] > c.hi || live4[r][k] < c.lo);
            if (valid[r] && (!UNAL || x0 + k < g.nx) &&
                (MODE == KF_MODE_STEP ? (moved || in_band) : (!quiet || in_band)))
                need |= 1u << (r * VEC + k);
        }
    }

    // Defaults of the voxels that cannot contribute, four at a time; phase-2 voxels overwrite
    // their own elements below (same thread, so program order holds).
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        if (!valid[r])
            continue;
        const unsigned int i0 = g.idx32(x0, yr[r], z);
        if (MODE == KF_MODE_ENERGY) {
            if (e_voxel)
                put_group<UNAL>(e_voxel + i0, make_float4(0.f, 0.f, 0.f, 0.f), g.nx - x0);
        } else {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float4 d = MODE == KF_MODE_GRAD ? make_float4(0.f, 0.f, 0.f, 0.f) : C[r][a];
                put_group<UNAL>(u_out + ((unsigned int)a * n32 + i0), d, g.nx - x0);
            }
        }
    }

    // Phase 2 -- the voxels that can contribute: full energy and gradient.  The wavefront's
    // need bits become one job list (ordered by bit, then lane), so a band segment held by one
    // or two lanes is spread over the wavefront: a lane runs one gather chain instead of up to
    // eight in a row.  A job's stores may hit elements another lane of this wavefront stored
    // as defaults above; a wavefront's vector memory operations are performed in order.
    unsigned long long bal[RPT * VEC];
    int total = 0;
#pragma unroll
    for (int b = 0; b < RPT * VEC; b++) {
        bal[b] = __ballot((need >> b) & 1u);
        total += __popcll(bal[b]);
    }
    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;
    for (int j = lane; j < total; j += 64) {
        int bit, src;
        kf_job_decode<RPT * VEC>(bal, j, bit, src);
        const int x = seg * ROW + src * VEC + bit % VEC;
        const int y = yt * ROWS + w + (bit / VEC) * (BLOCK / 64);
        if (x < g.nx && y < g.ny)
            adj_job<S, MODE>(L, Cn, u_in, u_out, c, &st->div_slot[parity], e_voxel, x, y, z, mx, esum, act);
    }
    if (MODE != KF_MODE_GRAD)
        kf_block_publish<BLOCK>(partials, blockIdx.x, esum, mx, act);
}

// ---------------------------------------------------------------------------
// Jacobi step over the flat voxel index: a thread owns FPT groups of 4 consecutive voxels (one
// float4 of each u plane each, FB VEC voxels apart), a block FLAT consecutive voxels, whatever
// the row length -- a step's streaming part is the same for every voxel, so rows that are not a
// multiple of the wavefront's width leave no lanes idle.  (x, y, z) is recovered only for the
// compacted stencil jobs.
// ---------------------------------------------------------------------------
constexpr int FPT = 2;                 // float4 groups per thread
constexpr int FB = BLOCK;              // threads per block
constexpr int FLAT = FB * VEC * FPT;   // voxels per block

// UNAL: N is not a multiple of 4, so the u planes a n (a = 1, 2) start off a 16-byte boundary.
// Their float4 groups are then read and written as 4-byte-aligned 16-byte accesses (one load
// or store each on this target), and the one group that runs past N goes element by element.
// (A 255^3 grid, with a store per voxel: 73 us per launch, against 59 us at 256^3, 108 us with
// one voxel per thread and 256 us for k_adj_any's full stencil per voxel.)
//
// Stores.  The destination buffer of launch k already holds u_{k-1}: launch k - 1 read it as
// its source, and every voxel outside the need set of launch k - 1 (active, or already moved)
// kept its value there, so u_k = u_{k-1} at those voxels.  A thread therefore stores its
// group's copy only when one of its voxels was in launch k - 1's need set (need words: one
// ballot per voxel of each group, FPT VEC 64-bit words per wavefront, written by the same
// wavefront of the previous launch), or when `full` is set (the first launch of a deform call,
// whose destination holds anything); active voxels are stored by the stencil below as before.
// Every launch leaves its destination equal to u_{k+1} at every voxel, bit for bit as with a
// store per voxel.  (256^3 BF16, one group per thread: 44.8 us per launch, against 57 us with
// a store per voxel; the launch then streams u and phi_live once and writes only the band's
// groups.)
//
// WIDE: 3 N >= 2^30, so plane offsets and voxel indices are 64-bit and the stencils run
// adj_job_wide; everything else is the same.  EXACT: N is a multiple of FLAT, so every group of
// every block lies inside the grid and the loads and stores go without bounds checks (256^3
// BF16: 42.0-42.3 against 43.0-43.8 us).
template <class S, bool UNAL, bool WIDE, bool EXACT>
__global__ __launch_bounds__(FB) void k_adj_flat(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                    BlockPartial *partials, int parity, const BlockPartial *prev,
                                                    unsigned int prev_nb, unsigned long long *__restrict__ needw,
                                                    int full)
{
    if (!step_gate<KF_MODE_STEP, FB>(st, prev, prev_nb, parity))
        return;
    using Ix = typename std::conditional<WIDE, unsigned long long, unsigned int>::type;
    const GridDims g = L.g;
    const Ix n32 = (Ix)g.n;
    // Tiles go to the XCDs in chunks of 64 (a z-slice at 256^3) dealt round-robin, so the
    // stencils of the surface band spread over every XCD instead of keeping the two or three
    // that hold the band busy after the others finish (256^3: 70 -> 58 us per launch).  With
    // this dealing every launch walks the whole grid at once, so sweeping odd launches
    // backwards (to meet the previous launch's last writes in the last-level cache) no longer
    // pays: it measured 58.9 against 58.1 us at 256^3.
    const unsigned int tile = kf_xcd_chunk_tile<64>(blockIdx.x, gridDim.x);
    // this wavefront's need words (FPT VEC of them: word q VEC + k = voxel k of each lane's group q)
    constexpr int NWW = FPT * VEC;
    const size_t nw = ((size_t)tile * (FB / 64) + (threadIdx.x >> 6)) * NWW;
    bool store_q[FPT];
#pragma unroll
    for (int q = 0; q < FPT; q++) {
        store_q[q] = full != 0;
        if (!full) {
            const ulonglong2 w01 = *reinterpret_cast<const ulonglong2 *>(needw + nw + q * 4);
            const ulonglong2 w23 = *reinterpret_cast<const ulonglong2 *>(needw + nw + q * 4 + 2);
            store_q[q] = (((w01.x | w01.y | w23.x | w23.y) >> (threadIdx.x & 63)) & 1ull) != 0ull;
        }
    }
    // group q of thread t: voxels tile FLAT + q FB VEC + t VEC .. + 3 (whole or outside unless
    // UNAL, where the last one may be partial)
    Ix i0[FPT];
    bool valid[FPT];
    float4 C[FPT][3];
    float live4[FPT][VEC];
    unsigned int nin[FPT];  // voxels of the group inside the grid
#pragma unroll
    for (int q = 0; q < FPT; q++) {
        i0[q] = (Ix)tile * (Ix)FLAT + (Ix)(q * (FB * VEC)) + (Ix)(threadIdx.x * VEC);
        valid[q] = EXACT || i0[q] < n32;
        nin[q] = !valid[q] ? 0u
                 : (UNAL && n32 - i0[q] < (Ix)VEC ? (unsigned int)(n32 - i0[q]) : (unsigned int)VEC);
#pragma unroll
        for (int k = 0; k < VEC; k++)
            live4[q][k] = 0.f;
        if (!UNAL || nin[q] == (unsigned int)VEC) {
#pragma unroll
            for (int a = 0; a < 3; a++)
                C[q][a] = !valid[q] ? make_float4(0.f, 0.f, 0.f, 0.f)
                          : UNAL    ? kf_load_stream4u(u_in + ((Ix)a * n32 + i0[q]))
                                    : kf_load_stream4(u_in + ((Ix)a * n32 + i0[q]));
            if (valid[q])
                S::load4(L.p, (int64_t)i0[q], L.scale, live4[q]);
        } else {
            float t[3][VEC];
#pragma unroll
            for (int k = 0; k < VEC; k++) {
#pragma unroll
                for (int a = 0; a < 3; a++)
                    t[a][k] = (unsigned int)k < nin[q] ? u_in[(Ix)a * n32 + i0[q] + (Ix)k] : 0.f;
                if ((unsigned int)k < nin[q])
                    live4[q][k] = S::load(L.p, (int64_t)(i0[q] + (Ix)k), L.scale);
            }
#pragma unroll
            for (int a = 0; a < 3; a++)
                C[q][a] = make_float4(t[a][0], t[a][1], t[a][2], t[a][3]);
        }
    }
    unsigned int need = 0u;
#pragma unroll
    for (int q = 0; q < FPT; q++) {
        if (!valid[q])
            continue;
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            if (UNAL && (unsigned int)k >= nin[q])
                continue;
            const bool moved = ((__float_as_uint(f4(C[q][0], k)) | __float_as_uint(f4(C[q][1], k)) |
                                 __float_as_uint(f4(C[q][2], k))) & 0x7fffffffu) != 0u;
            const bool in_band = !(live4[q][k] > c.hi || live4[q][k] < c.lo);
            if (moved || in_band)
                need |= 1u << (q * VEC + k);
        }
    }
    // The stencil for active or moved voxels, compacted over the wavefront as in k_adj_rows.
    unsigned long long bal[FPT * VEC];
    int total = 0;
#pragma unroll
    for (int b = 0; b < FPT * VEC; b++) {
        bal[b] = __ballot((need >> b) & 1u);
        total += __popcll(bal[b]);
    }
#pragma unroll
    for (int q = 0; q < FPT; q++) {
        if (!valid[q] || !store_q[q])
            continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float *dst = u_out + ((Ix)a * n32 + i0[q]);
            if (!UNAL) {
                *reinterpret_cast<float4 *>(dst) = C[q][a];
            } else if (nin[q] == (unsigned int)VEC) {
                *reinterpret_cast<kf_f4u *>(dst) = kf_f4u{C[q][a].x, C[q][a].y, C[q][a].z, C[q][a].w};
            } else {
#pragma unroll
                for (int k = 0; k < VEC; k++)
                    if ((unsigned int)k < nin[q])
                        dst[k] = f4(C[q][a], k);
            }
        }
    }
    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;
    const Ix wave0 = (Ix)tile * (Ix)FLAT + (Ix)((threadIdx.x & ~63u) * VEC);
    for (int j = threadIdx.x & 63; j < total; j += 64) {
        int bit, src;
        kf_job_decode<FPT * VEC>(bal, j, bit, src);
        const Ix i = wave0 + (Ix)((bit / VEC) * (FB * VEC) + src * VEC + bit % VEC);
        int x, y, z;
        if (WIDE) {
            kf_xyz(g, (int64_t)i, x, y, z);
            adj_job_wide<S, KF_MODE_STEP, false>(L, Cn, u_in, u_out, c, &st->div_slot[parity], nullptr,
                                                 (int64_t)i, x, y, z, mx, esum, act);
        } else {
            g.xyz32((unsigned int)i, x, y, z);
            adj_job<S, KF_MODE_STEP>(L, Cn, u_in, u_out, c, &st->div_slot[parity], nullptr, x, y, z, mx, esum,
                                     act);
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int b = 0; b < NWW; b += 2)
            *reinterpret_cast<ulonglong2 *>(needw + nw + b) = make_ulonglong2(bal[b], bal[b + 1]);
    }
    kf_block_publish<FB>(partials, blockIdx.x, esum, mx, act, total > 0);
}

// Nonzero mask of field values u (layout above; UNAL as in k_adj_rows).
template <bool UNAL>
__global__ __launch_bounds__(BLOCK) void k_adj_mask(const float *__restrict__ u, GridDims g, int nseg, int nyt,
                                                    unsigned int *__restrict__ mask)
{
    const unsigned int tile = blockIdx.x;
    const int seg = (int)(tile % (unsigned)nseg);
    const int yt = (int)((tile / (unsigned)nseg) % (unsigned)nyt);
    const int z = (int)(tile / ((unsigned)nseg * (unsigned)nyt));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x0 = seg * ROW + lane * VEC;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int y = yt * ROWS + w + r * (BLOCK / 64);
        unsigned int b = 0u;
#pragma unroll
        for (int a = 0; a < 3; a++)
            b |= bits4(row_group<UNAL>(u, g, a, x0, y, z));
        put_bits(mask, nseg * MW, g, y, z, seg, lane, (b & 0x7fffffffu) != 0u);
    }
}

template <class S, int MODE>
void launch(const PhiRead<S> &L, const PhiRead<S> &Cn, const float *u_in, float *u_out, const AdjConsts &c,
            kf_context *ctx, float *e_voxel, bool rows, bool flat1, bool wide, unsigned int nb, int nseg, int nyt,
            const unsigned int *mask_in, int parity, const BlockPartial *prev, unsigned int prev_nb, int full)
{
    BlockPartial *parts = kf_partials(ctx, parity);
    const bool exact = L.g.n % FLAT == 0;
    if (MODE == KF_MODE_STEP && rows && exact)
        hipLaunchKernelGGL((k_adj_flat<S, false, false, true>), dim3(nb), dim3(FB), 0, ctx->stream, L, Cn, u_in,
                           u_out, c, ctx->state, parts, parity, prev, prev_nb, ctx->need_words, full);
    else if (MODE == KF_MODE_STEP && rows)
        hipLaunchKernelGGL((k_adj_flat<S, false, false, false>), dim3(nb), dim3(FB), 0, ctx->stream, L, Cn, u_in,
                           u_out, c, ctx->state, parts, parity, prev, prev_nb, ctx->need_words, full);
    else if (MODE == KF_MODE_STEP && flat1)
        hipLaunchKernelGGL((k_adj_flat<S, true, false, false>), dim3(nb), dim3(FB), 0, ctx->stream, L, Cn, u_in,
                           u_out, c, ctx->state, parts, parity, prev, prev_nb, ctx->need_words, full);
    else if (MODE == KF_MODE_STEP && wide && exact)
        hipLaunchKernelGGL((k_adj_flat<S, false, true, true>), dim3(nb), dim3(FB), 0, ctx->stream, L, Cn, u_in,
                           u_out, c, ctx->state, parts, parity, prev, prev_nb, ctx->need_words, full);
    else if (MODE == KF_MODE_STEP && wide && L.g.n % VEC == 0)
        hipLaunchKernelGGL((k_adj_flat<S, false, true, false>), dim3(nb), dim3(FB), 0, ctx->stream, L, Cn, u_in,
                           u_out, c, ctx->state, parts, parity, prev, prev_nb, ctx->need_words, full);
    else if (MODE == KF_MODE_STEP && wide)
        hipLaunchKernelGGL((k_adj_flat<S, true, true, false>), dim3(nb), dim3(FB), 0, ctx->stream, L, Cn, u_in,
                           u_out, c, ctx->state, parts, parity, prev, prev_nb, ctx->need_words, full);
    else if (rows && L.g.nx % VEC == 0)
        hipLaunchKernelGGL((k_adj_rows<S, MODE, false>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in, u_out,
                           c, ctx->state, parts, e_voxel, nseg, nyt, mask_in, parity, prev, prev_nb);
    else if (rows)
        hipLaunchKernelGGL((k_adj_rows<S, MODE, true>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in, u_out,
                           c, ctx->state, parts, e_voxel, nseg, nyt, mask_in, parity, prev, prev_nb);
    else
        hipLaunchKernelGGL((k_adj_any<S, MODE>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in, u_out, c,
                           ctx->state, parts, e_voxel, parity, prev, prev_nb);
}

}  // namespace

size_t kf_adj_mask_words(const int32_t dims[3])
{
    if (3 * (int64_t)dims[0] * dims[1] * dims[2] >= ((int64_t)1 << 30))
        return 0;
    return (size_t)((dims[0] + ROW - 1) / ROW) * MW * (size_t)dims[1] * (size_t)dims[2];
}

kf_status kf_launch_adj_mask(const kf_field *f, const float *u, unsigned int *mask)
{
    const int nseg = (f->dims[0] + ROW - 1) / ROW, nyt = (f->dims[1] + ROWS - 1) / ROWS;
    const size_t words = kf_adj_mask_words(f->dims);
    KF_CHECK_ARG(words > 0, "nonzero masks need 3 N < 2^30 (32-bit offsets)");
    const unsigned int nb = (unsigned int)(nseg * nyt * f->dims[2]);
    if (f->dims[0] % VEC == 0)
        hipLaunchKernelGGL(k_adj_mask<false>, dim3(nb), dim3(BLOCK), 0, f->ctx->stream, u, grid_of(f), nseg, nyt,
                           mask + words);
    else
        hipLaunchKernelGGL(k_adj_mask<true>, dim3(nb), dim3(BLOCK), 0, f->ctx->stream, u, grid_of(f), nseg, nyt,
                           mask + words);
    hipLaunchKernelGGL(k_adj_dilate, dim3((unsigned int)((words + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                       f->ctx->stream, mask + words, grid_of(f), nseg * MW, mask);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_adj_step(const kf_volume *live, const kf_volume *can, const float *u_in, float *u_out,
                             const kf_params *p, int mode, float *e_voxel, const unsigned int *mask_in, int iter,
                             unsigned int *nb_out)
{
    kf_context *ctx = live->ctx;
    const AdjConsts c = kf_adj_consts(p);
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    // A step runs the flat kernel when N is a multiple of 4 and 3 N < 2^30 (float4 plane
    // accesses, 32-bit offsets), its unaligned form for other N < 2^30 / 3, and its 64-bit
    // forms (aligned or not) on larger grids; energy and gradient passes run the row kernel
    // (its unaligned form for rows that are not a multiple of 4 voxels) when 3 N < 2^30 and the
    // nonzero mask of u_in is given.  Everything else runs k_adj_any (a full stencil per voxel).
    const bool small = 3 * live->n < ((int64_t)1 << 30);
    const bool flat = mode == KF_MODE_STEP && live->n % VEC == 0 && small;
    const bool flat1 = mode == KF_MODE_STEP && !flat && small;
    const bool wide = mode == KF_MODE_STEP && !small;
    const bool rows = flat || (mode != KF_MODE_STEP && kf_adj_mask_words(live->dims) > 0 && mask_in != nullptr);
    const int nseg = (live->dims[0] + ROW - 1) / ROW;
    const int nyt = (live->dims[1] + ROWS - 1) / ROWS;
    const unsigned int nb = flat || flat1 || wide ? (unsigned int)((live->n + FLAT - 1) / FLAT)
                            : rows                ? (unsigned int)(nseg * nyt * live->dims[2])
                                                  : (unsigned int)((live->n + BLOCK - 1) / BLOCK);
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s == KF_OK && (flat || flat1 || wide))
        s = kf_ensure_need_words(ctx, (size_t)nb * (FB / 64) * (FPT * VEC));  // a word per ballot
    if (s != KF_OK)
        return s;
    if (nb_out)
        *nb_out = nb;
    // Energy and gradient passes use partial set 0; step launch k uses set k & 1 and folds
    // launch k - 1 (same shape: the dims do not change within a deform call).
    const int parity = mode == KF_MODE_STEP ? (iter & 1) : 0;
    const BlockPartial *prev = kf_partials(ctx, parity ^ 1);
    const unsigned int prev_nb = mode == KF_MODE_STEP && iter > 0 ? nb : 0u;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        PhiRead<S> L = make_read<S>(live), Cn = make_read<S>(can);
        if (mode == KF_MODE_GRAD)
            launch<S, KF_MODE_GRAD>(L, Cn, u_in, u_out, c, ctx, e_voxel, rows, false, false, nb, nseg, nyt, mask_in,
                                    0, prev, 0u, 1);
        else if (mode == KF_MODE_ENERGY)
            launch<S, KF_MODE_ENERGY>(L, Cn, u_in, u_out, c, ctx, e_voxel, rows, false, false, nb, nseg, nyt,
                                      mask_in, 0, prev, 0u, 1);
        else  // the first step of a deform call stores every voxel
            launch<S, KF_MODE_STEP>(L, Cn, u_in, u_out, c, ctx, e_voxel, rows, flat1, wide, nb, nseg, nyt, nullptr,
                                    parity, prev, prev_nb, iter == 0 ? 1 : 0);
    });
    KF_LAUNCH_CHECK();
    return mode == KF_MODE_ENERGY ? kf_launch_finalize(ctx, nb, 0) : KF_OK;
}
