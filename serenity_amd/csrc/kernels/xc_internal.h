// Device-side data layout of the XC Fock build and the kernel launchers shared
// by the C ABI (xc_api.hip) and the kernels (xc_kernels.hip).
//
// HBM layout of the basis-functions-on-grid tiles: the grid is cut into
// blocks of at most 128 consecutive (Hilbert-ordered) points.  Block b keeps
// the n_b AOs that survive Serenity's shell screening, padded to npad (a
// multiple of 32) and stored as ncomp component planes
//   tiles[tile_off + (c * npad + i) * 128 + p],  c = value, d/dx, d/dy, d/dz
// (and, for highest derivative 2, d2/dxx, dxy, dxz, dyy, dyz, dzz),
// i = compact AO index (ao_idx maps it to the global AO, -1 for padding),
// p = point within the block (rows beyond the block's size are zero).  Point
// is the fastest index so a wavefront reads 64 consecutive values.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "functional.h"

namespace sxc {

// Global-memory view of a pointer.  A pointer the compiler cannot trace to a kernel argument (one
// taken from a packed build's per-system table, or selected between such a pointer and an
// argument) is generic to it, and accesses through a generic pointer become flat instructions:
// those count against the LDS counter as well as the vector-memory one, so every wait for an LDS
// read then also waits for all global loads in flight.  Accesses through gmem<T>* are global ones.
template <typename T>
using gmem = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ gmem<T>* to_global(T* p) {
  return (gmem<T>*)p;
}
// Vector types for 16-byte global loads through gmem pointers (HIP's float4 and the like are
// classes whose copy constructors take generic references).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
// A struct read through a gmem pointer (structs, too, copy only from generic references), in
// 8-byte words.
template <typename S>
__device__ __forceinline__ S load_global(const gmem<const S>* p) {
  static_assert(sizeof(S) % 8 == 0 && alignof(S) >= 8, "8-byte words");
  S out;
  const gmem<const int64_t>* src = reinterpret_cast<const gmem<const int64_t>*>(p);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(S) / 8); ++k) {
    const int64_t w = src[k];
    __builtin_memcpy(reinterpret_cast<char*>(&out) + 8 * k, &w, 8);
  }
  return out;
}

constexpr int kTile = 128;      // points per tile row (max block size)
constexpr int kAoPad = 32;      // AO padding granularity of a block
constexpr int kMaxL = 6;        // highest angular momentum the AO kernel evaluates

struct DevShell {
  double cx, cy, cz;
  int l, pure, nprim, prim_off;
  int ao_off, nfunc, ang_off, ncart;
};

struct DevBlock {
  int64_t tile_off;   // element offset of the block's tiles
  int64_t toff;       // element offset of the block's Fock sub-tiles in the sub-tile scratch
  int64_t first;      // first grid point
  int npts;           // points in the block (<= 128)
  int nb;             // non-negligible AOs
  int npad;           // nb rounded up to kAoPad
  int idx_off;        // offset into ao_idx
  int shell_off;      // offset into the significant-shell list
  int nshell;         // significant shells
  int npass;          // density column passes (partial rho/grad slots written)
  int sys;            // system of a packed build (index into its PackSys table; 0 otherwise)
};

// Fock sub-tiles of a block: the npad / 16 = nt sub-tile rows and columns of its upper
// triangle, (I, J) with I <= J, in the block's slice of the scratch (DevBlock::toff), stored by
// matrix rows: compact row a (in sub-tile row I = a / 16) holds the columns 16 I .. npad - 1 (from
// its diagonal sub-tile on) contiguously, and the 16 rows of sub-tile row I follow those of the
// rows above, starting at element 256 fock_subtile_index(I, I, nt).  The reduce then reads each
// AO row's segment of a block as one contiguous run.
__host__ __device__ inline int64_t fock_subtile_index(int I, int J, int nt) {
  return (int64_t)I * nt - (int64_t)I * (I - 1) / 2 + (J - I);
}
__host__ __device__ inline int64_t fock_subtile_elems(int npad) {
  const int64_t nt = npad / 16;
  return nt * (nt + 1) / 2 * 256;
}
// element of compact row a, column 16 (a / 16), in the block's slice
__host__ __device__ inline int64_t fock_row_start(int a, int npad) {
  const int I = a / 16;
  return 256 * fock_subtile_index(I, I, npad / 16) + (int64_t)(a % 16) * (npad - 16 * I);
}

// One contribution to a row of the Fock matrix: block `block` holds AO `row` at compact position
// c0, and the compact columns c0 .. c1 - 1 (AOs idx[idx_off + c], in increasing global order)
// carry the values F[row][idx[c]] of the block, column c at element rowbase + c of the scratch
// (rowbase = the block's toff + fock_row_start(c0, npad) - 16 (c0 / 16)).
struct RowEnt {
  int64_t rowbase;
  int c0, c1, idx_off, block;
};

// One system of a packed build: several controllers' builds in one launch per
// stage.  Work items index a merged block list whose entries keep their own
// system's offsets (tile_off, first, idx_off) and name the system in DevBlock::sys.
struct PackSys {
  const void* tiles;    // AO tiles in the contraction precision
  const void* P;        // density matrix in the contraction precision
  const double* P64;    // density matrix (FP64)
  float* P32;           // FP32 copy (FP32 builds)
  const int* ao_idx;
  double* part;         // partial densities
  const double* w;      // grid weights
  double* vs;           // w v
  double* gw;           // w g
  double* U;            // Fock accumulator (lower triangle, written by the row reduce)
  void* T;              // Fock sub-tile scratch (contraction precision)
  const RowEnt* ent;    // row entries of the reduce (sorted by row, then block)
  const int* row_ptr;   // n + 1 offsets into ent
  double* F;            // XC matrix out
  double* eblk;         // per-block energies
  double* round;        // rounding-term partials (FP32)
  double* E;            // E_xc out (device)
  int64_t npts;
  int n;
  int blk0;             // first merged block of this system
  int nxc;              // 128-point tiles of its grid
  int pad_;
};

// Work item of the sub-tile Fock kernel: the block's AO rows [a0, a0 + na)
// (A side) and [b0, b0 + nb) (B side), multiples of 16.  nb == 0 marks a
// diagonal item (B = A, sub-tiles i <= j).  nsub = 16 x 16 sub-tiles.
struct FockItem {
  int block;
  int a0, na;
  int b0, nb;
  int nsub;
};
constexpr int kFockRows = 256;        // rows a sub-tile item may stage
constexpr int kFockPanel = 128;       // panel height of blocks split into panel pairs
#ifndef SXC_FOCK_WAVES
#define SXC_FOCK_WAVES 16
#endif
constexpr int kFockWaves = SXC_FOCK_WAVES;  // wavefronts per workgroup of the sub-tile kernel
#if SXC_FOCK_WAVES == 16
constexpr int kFockNumS = 5;          // slot-count classes of the sub-tile kernel
constexpr int kFockS[kFockNumS] = {1, 2, 3, 4, 5};
inline int fock_smax(int) { return 5; }  // accumulator budget: kFockWaves * 5 sub-tiles per item
#else
constexpr int kFockNumS = 9;
constexpr int kFockS[kFockNumS] = {1, 2, 3, 4, 5, 6, 7, 8, 10};
inline int fock_smax(int) { return 10; }
#endif

// Work item of the density kernel: block, a = first column of the pass,
// b = columns of the pass (multiple of 16), pass = partial-density slot.
struct WorkItem {
  int block, a, b, pass;
};

// ---- launchers (xc_kernels.hip) -------------------------------------------
void launch_screen(hipStream_t s, const double* pts, int64_t npts, int bs, int nblocks, const DevShell* shells,
                   int nshell, const double* alpha, const double* coef, double radial_thr, unsigned char* flags);

void launch_ao_eval(hipStream_t s, const DevBlock* blocks, int nblocks, const int* sig_shell, const int* sig_pos,
                    const DevShell* shells, const double* alpha, const double* coef, const double* ang,
                    const double* pts, double radial_thr, double exp_thr, int ncomp, double* tiles64, float* tiles32);

// Sub-tile density kernel: 8 wavefronts, S = pass columns / 16 (1..kDensMaxS).
// Partial densities: part[(pass * 4 + c) * npts + q], c = rho, d/dx, d/dy, d/dz
// (gradient slots hold sum T grad phi without the factor 2).
constexpr int kDensMaxS = 6;
// Mixed precision: blocks whose largest density (bohr^-3) after the FP32 density pass exceeds
// this are contracted again in FP64 (the cores of the heavy atoms, where FP32 rounding of the AO
// values and of phi P gives a bias that does not cancel over the grid).
constexpr double kMixedRhoThreshold = 2.0;
constexpr int kDensIdxLds = 2048;  // AO index lists up to this length are staged in LDS
// With nwork_dev, the launch covers a list filled on the device: only its first *nwork_dev
// items run (nwork is the list's capacity).
template <typename T>
void launch_density_sub(hipStream_t s, int S, const WorkItem* work, int nwork, const DevBlock* blocks,
                        const int* ao_idx, const T* tiles, const T* P, int n, int want_grad, double* part,
                        int64_t npts, const PackSys* pack = nullptr, const int* nwork_dev = nullptr,
                        const unsigned char* skip = nullptr);

// Dense-block refinement of the mixed-precision density: the work lists (one per column-tile
// class) and their device-side counts, filled by launch_mark_dense.
struct DenseLists {
  WorkItem* item[kDensMaxS + 1];
  int* count;  // kDensMaxS + 1 counters
};
// Clears the counts, then appends the density items of every block whose largest density in the
// partials part (one pass per item, 128-point blocks) exceeds rho_thr; by_block holds every block's
// items contiguously from item0[b] (blocks[b].npass of them), cls[b] is their class.
// With a mask (one byte per block: the blocks the first pass skipped because the previous build
// found them dense), only the blocks outside it are tested and the listed ones join it;
// launch_list_masked lists the mask's own blocks (before the first pass, so that their second pass
// can run beside it), and launch_refresh_dense, after the second passes, keeps in the mask only the
// blocks whose density from them still exceeds rho_thr.
void launch_mark_dense(hipStream_t s, const DevBlock* blocks, int nblocks, int64_t npts, const double* part,
                       double rho_thr, const WorkItem* by_block, const int* item0, const int* cls,
                       const DenseLists& lists, unsigned char* mask = nullptr);
void launch_list_masked(hipStream_t s, const DevBlock* blocks, int nblocks, const WorkItem* by_block,
                        const int* item0, const int* cls, const DenseLists& lists, const unsigned char* mask);
void launch_refresh_dense(hipStream_t s, const DevBlock* blocks, int nblocks, int64_t npts, const double* part,
                          double rho_thr, unsigned char* mask);
void launch_to_double(hipStream_t s, const float* src, double* dst, int64_t n);

// rho[q] / grad[.][q] from the partials of the block of q (any block size).
void launch_reduce_partials(hipStream_t s, const DevBlock* blocks, int bs, int64_t npts, const double* part,
                            int want_grad, double* rho, double* grad, int accumulate = 0);

// Functional on rho/grad, or (part != nullptr) on the partials of 128-point blocks.
void launch_functional(hipStream_t s, int func, const XcParams& prm, int64_t npts, int bs, const double* rho,
                       const double* grad, const DevBlock* blocks, const double* part, const double* w, double* eps,
                       double* v, double* g, double* vs, double* gw, double* e_blk,
                       const double* rho_sub = nullptr, const double* grad_sub = nullptr);
// Functional of a packed build: one 128-point block per workgroup over the merged block list.
void launch_functional_packed(hipStream_t s, int func, const XcParams& prm, int nblocks, const DevBlock* blocks,
                              const PackSys* pack);

// Unrestricted functional: spin densities/gradients (3 x npts each) in; w v_s, w g_s and
// block energies out.
void launch_functional_pol(hipStream_t s, int func, const XcParams& prm, int64_t npts, int bs, const double* ra,
                           const double* ga, const double* rb, const double* gb, const double* w, double* vsa,
                           double* gwa, double* vsb, double* gwb, double* e_blk,
                           const double* ra2 = nullptr, const double* ga2 = nullptr, const double* rb2 = nullptr,
                           const double* gb2 = nullptr);

// Sub-tile Fock kernel: kFockWaves wavefronts, slot class sc (kFockS[sc] sub-tiles per wavefront).
// Each item stores its sub-tiles of phi^T G + G^T phi into the block's scratch slice Tscr + toff
// (zeros when the block-average test skips the block).
template <typename T>
void launch_fock_sub(hipStream_t s, int sc, const FockItem* work, int nwork, const DevBlock* blocks,
                     const int* ao_idx, const T* tiles, const double* vs, const double* gw, int gga,
                     double block_ave_thr, T* Tscr, int64_t npts, const PackSys* pack = nullptr);
// U[j + n i] = sum over the row entries of AO i (block order) of the stored sub-tile values at
// columns j >= i, for every i of [0, n): the lower triangle of the Fock accumulator, each element
// written once (its terms meet in LDS in an order that varies from run to run).  Packed:
// blockIdx.y = system, n = the largest n.  16 wavefronts per row: at (H2O)64/def2-TZVP the reduce
// takes 0.68 ms against 0.86 ms with 8 and 1.40 ms with 4 (each wavefront's entries form a chain
// of dependent load-add steps; more wavefronts, shorter chains).
constexpr int kReduceWaves = 16;
constexpr int kReduceWindow = 8192;  // row columns accumulated in LDS per pass
template <typename T>
void launch_fock_reduce(hipStream_t s, const RowEnt* ent, const int* row_ptr, const T* Tscr, const int* ao_idx,
                        double* U, int n, const PackSys* pack = nullptr, int nsys = 1, bool fixed = false);
// Packed elementwise stages over nsys systems (maxnn = largest n^2): P -> FP32,
// F = mirror of U (+ FP32 rounding partials), E = sum of block energies (+ rounding term).
constexpr int kPackRoundBlocks = 512;  // finalize workgroups per system of a packed build
void launch_pack_to_float(hipStream_t s, const PackSys* pack, int nsys, int64_t maxnn);
void launch_pack_finalize(hipStream_t s, const PackSys* pack, int nsys, int fp32);
void launch_pack_sum(hipStream_t s, const PackSys* pack, int nsys, int fp32);

// Work item of the gradient kernel: the block's AO columns [col0, col0 + 64).
struct GradItem {
  int block, col0;
};
constexpr int kGradCols = 64;
// a[nu][c] += sum over the item's points of (G P)_p,nu d_c phi_nu + (phi P)_p,nu H^c_nu
// (kernels/gradient_kernels.hip); tiles need 4 planes (LDA) or 10 (GGA).
void launch_gradient(hipStream_t s, const GradItem* work, int nwork, const DevBlock* blocks, const int* ao_idx,
                     const double* tiles, const double* P, int n, const double* vs, const double* gw, int gga,
                     int64_t npts, double* acc);
// grad[A + natoms c] = -2 sum_{nu on A} acc[nu][c].
void launch_gradient_atoms(hipStream_t s, const double* acc, const int* ao_atom, int n, int natoms, double* grad);

// ---- density-fitted Coulomb matrix (kernels/coulomb_kernels.hip) ----------
// One significant primitive pair (a of shell A, b of shell B, A >= B) of the orbital basis:
// p = a + b, K = c_a c_b exp(-a b / p |AB|^2), centre P, P - A, P - B; sym = 2 for A > B
// (the pair stands for both orders), 1 for A == B; herm_off: offset of its (L + 1)(L + 2)(L + 3) / 6
// Hermite coefficients (L = l_A + l_B) in the per-pair arrays.
struct JPair {
  int A, B, L, herm_off;
  double sym, p, K;
  double P[3], PA[3], PB[3];
};
constexpr int kJMaxLPrim = 3;                      // orbital shells up to f
constexpr int kJMaxLAux = 4;                       // auxiliary shells up to g
constexpr int kJMaxL = 2 * kJMaxLPrim + kJMaxLAux;
constexpr int kJMaxH = (kJMaxL + 1) * (kJMaxL + 2) * (kJMaxL + 3) / 6;
constexpr int kJMaxHab = (2 * kJMaxLPrim + 1) * (2 * kJMaxLPrim + 2) * (2 * kJMaxLPrim + 3) / 6;
constexpr int kJMaxHs = (kJMaxLAux + 1) * (kJMaxLAux + 2) * (kJMaxLAux + 3) / 6;
constexpr int kJMaxNcart = (kJMaxLPrim + 1) * (kJMaxLPrim + 2) / 2;
constexpr int kJMaxNcartAux = (kJMaxLAux + 1) * (kJMaxLAux + 2) / 2;
constexpr int kJMaxE = (kJMaxLPrim + 1) * (kJMaxLPrim + 1) * (2 * kJMaxLPrim + 1);
void launch_jdens(hipStream_t s, const JPair* pairs, int npair, const DevShell* shells, const double* ang,
                  const double* P, int n, double* D);
void launch_jrhs(hipStream_t s, const JPair* pairs, int npair, const double* D, const DevShell* aux, int naux_shell,
                 const double* aux_alpha, const double* aux_coef, const int* aux_cart_off, double* bc);
void launch_jpot(hipStream_t s, const JPair* pairs, int npair, const DevShell* aux, int naux_shell,
                 const double* aux_alpha, const double* aux_coef, const int* aux_cart_off, const double* cc,
                 double* V);
void launch_jassemble(hipStream_t s, const JPair* pairs, int npair, const DevShell* shells, const double* ang,
                      const double* V, int n, double* J);

// Molecular-grid partition weights of ncand candidate points (kernels/grid_kernels.hip).
void launch_partition(hipStream_t s, int type, unsigned smoothing, int natoms, const double* xyz, const double* dist,
                      const double* apair, const double* min_dist, int64_t ncand, const double* cand_xyz,
                      const double* cand_w, const int* cand_atom, double* w);

// The device grid build (kernels/grid_kernels.hip): partition weights of every candidate generated
// from the reference grids (keep[c] = w[c] > thr), the cut (exclusive scan of keep into pos, the
// kept points in candidate order), the bounding box (mn, spread) and the Hilbert order.
void launch_partition_gen(hipStream_t s, int type, unsigned smoothing, int natoms, const double* xyz,
                          const double* dist, const double* apair, const double* min_dist, int64_t ncand,
                          const double* ref_xyz, const double* ref_w, const int64_t* ref_off,
                          const int64_t* cand_off, double thr, double* w, int* keep);
size_t grid_scan_temp_bytes(int64_t n);
void launch_cut(hipStream_t s, int natoms, const double* xyz, int64_t ncand, const double* ref_xyz,
                const int64_t* ref_off, const int64_t* cand_off, const double* w, const int* keep, int* pos,
                void* temp, size_t temp_bytes, double* out_xyz, double* out_w);
int bbox_part_count();
void launch_bbox(hipStream_t s, const double* pts, int64_t n, int64_t length, double* part, double* box);
size_t grid_sort_temp_bytes(int64_t n);
void launch_hilbert_order(hipStream_t s, const double* pts, const double* w, int64_t n, const double* box,
                          int64_t length, int64_t* key, int64_t* key_sorted, int64_t* idx, int64_t* order,
                          void* temp, size_t temp_bytes, double* out_pts, double* out_w);

void launch_weight_potential(hipStream_t s, int64_t npts, const double* w, const double* v, const double* g,
                             double* vs, double* gw);
void launch_fill(hipStream_t s, double* p, int64_t n, double value);
void launch_copy(hipStream_t s, const void* src, void* dst, int64_t n32);  // n32 units of 32 bytes

// Up to kListMax device arrays passed by value (kernel argument).
constexpr int kListMax = 32;
struct DensityList {
  const double* rho[kListMax];
  const double* grad[kListMax];
  int n;
};
// rho (+)= sum of list.rho[i] in list order; grad likewise (3 x npts each).
void launch_sum_densities(hipStream_t s, const DensityList& list, int64_t npts, int want_grad, int accumulate,
                          double* rho, double* grad);
// out[0] -= list.rho[i][0] for i in order.
void launch_subtract_scalars(hipStream_t s, const DensityList& list, double* out);
void launch_to_float(hipStream_t s, const double* src, float* dst, int64_t n);
// F[i + n j] (+)= U[max(i,j) + n min(i,j)] (U holds the lower triangle)
// F = mirror of U (accumulate: F += it); with P32, part[b] = block b's sum of F (P - P32)
// (finalize_blocks(n) partials)
void launch_finalize(hipStream_t s, const double* U, double* F, int n, int accumulate, const double* P = nullptr,
                     const float* P32 = nullptr, double* part = nullptr);
int finalize_blocks(int n);
// *out = sum of x[0..n) and x2[0..n2) (subtract: *out -= that sum)
void launch_sum(hipStream_t s, const double* x, int64_t n, double* out, int subtract = 0, const double* x2 = nullptr,
                int64_t n2 = 0);

}  // namespace sxc
