// internal.h — host-side declarations shared by the translation units of
// libfmmh: device-buffer arena, per-stage launchers, the context.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fmmh.h"
#include "common.h"

namespace fmmh {

// ------------------------------------------------------------------ prims
size_t scan_tmp_words(size_t n);
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp,
                        hipStream_t s, uint32_t** total_dev);
size_t radix_tmp_words(size_t n);
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* k_alt, uint32_t* v_alt, size_t n,
                   int bit_lo, int bit_hi, uint32_t* tmp, hipStream_t s);
// Directed entries of undirected pairs (x, y): 2i -> (y << sbits | x), 2i+1
// -> (x << sbits | y), or with `sentinel` all ones in [0, 2 sbits) for a
// self pair's second entry; sorted stably by the target bits.
int radix_sort_directed(const uint2* pairs, uint32_t np, int sbits, bool sentinel, uint64_t* keys,
                        uint64_t* k_alt, uint32_t* tmp, hipStream_t s);

// Cells in DFS preorder (tree.rs:49-58 as structure of arrays).
struct DevCells {
  uint32_t* first;
  uint32_t* count;
  int32_t* children;  // [nc][8], -1 absent
  int32_t* parent;    // -1 for the root
  uint8_t* leaf;
  uint8_t* level;
  double* cx;
  double* cy;
  double* cz;
  double* mass;
  double* rho;
  int32_t* rope;        // next cell after this subtree when children run 7..0
  int32_t* last_child;  // highest-octant child, -1 for leaves
};

struct DevBodies {  // slot (Morton) order, bodies.rs:13-22
  double* x;
  double* y;
  double* z;
  double* m;
  double4* xyzm;  // the same bodies interleaved (one 32-byte load per body)
};

// BFS (level-order) scratch of the tree build.
struct DevBfs {
  uint32_t* first;
  uint32_t* count;
  int32_t* parent;  // BFS index of the parent
  uint8_t* oct;
  uint8_t* level;
  uint32_t* dfs;    // DFS id of each BFS cell
};

// -------------------------------------------------------------- tree.hip
void launch_bbox(const double* pos, size_t n, double* partial, BBox* out, hipStream_t s);
// Ensembles of nsys systems (body ranges off[s]..off[s+1], device array).
void launch_bbox_sys(const double* pos, const uint64_t* off, uint32_t nsys, BBox* out,
                     hipStream_t s);
void launch_keys_sys(const double* pos, const double* mass, size_t n, const uint64_t* off,
                     uint32_t nsys, int vlev, const BBox* bbs, uint64_t* keys, uint32_t* vals,
                     uint32_t* flags, double4* pk, hipStream_t s);
// Ensemble keys on bodies sorted by standalone key: system prefix over the
// key's top 63 - 3 vlev bits (then sorted stably on the prefix bits).
void launch_keys_prefix(uint64_t* keys, const uint32_t* ids, size_t n, const uint64_t* off,
                        uint32_t nsys, int vlev, hipStream_t s);
// Both keys kernels also write pk[i] = (x, y, z, m) of body i (id order).
void launch_keys(const double* pos, const double* mass, size_t n, const BBox* bb,
                 uint64_t* keys, uint32_t* vals, uint32_t* flags, double4* pk, hipStream_t s);
void launch_gather(const double* pos, const double* mass, const uint32_t* order, size_t n,
                   DevBodies b, hipStream_t s);
void launch_gather_packed(const double4* pk, const uint32_t* order, size_t n, DevBodies b,
                          hipStream_t s);
void launch_tree_init(DevBfs bfs, size_t n, uint32_t* chain, uint8_t* top, hipStream_t s);
void launch_split_count(DevBfs bfs, uint32_t s0, uint32_t cnt, const uint64_t* keys,
                        uint32_t n_leaf, int vlev, uint32_t* nchild, uint32_t* bounds,
                        hipStream_t s);
void launch_split_emit(DevBfs bfs, uint32_t s0, uint32_t cnt, uint32_t e0, uint32_t n_leaf,
                       int vlev, const uint32_t* bounds, const uint32_t* child_off,
                       uint32_t* chain, uint8_t* top, hipStream_t s);
void launch_dfs_ids(DevBfs bfs, uint32_t nc, const uint32_t* cstart, const uint8_t* top,
                    hipStream_t s);
void launch_scatter_cells(DevBfs bfs, uint32_t nc, uint32_t n_leaf, int vlev, DevCells c,
                          uint32_t* lvl_cells, hipStream_t s);
// which = 0: all cells of the list, 1: its leaves only, 2: its internal cells only.
void launch_geometry_leaves(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                            int center_mode, hipStream_t s);
void launch_geometry(const uint32_t* lvl_cells, uint32_t cnt, DevCells c, DevBodies b,
                     int center_mode, hipStream_t s, int which);
void launch_ropes(const uint32_t* lvl_cells, uint32_t cnt, DevCells c, hipStream_t s);
void launch_leaf_flags(DevCells c, uint32_t nc, uint32_t* flags, hipStream_t s);
void launch_leaves_by_count(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c,
                            uint64_t* keys, uint64_t* keys_alt, uint32_t* tmp, uint32_t* out,
                            hipStream_t s);
void launch_internal_flags(const uint32_t* lvl_cells, uint32_t nc, DevCells c, uint32_t* flags,
                           hipStream_t s);
void launch_internal_compact(const uint32_t* lvl_cells, uint32_t nc, DevCells c,
                             const uint32_t* pos, uint32_t* out, hipStream_t s);
void launch_level_pick(const uint32_t* pos, const uint32_t* total, uint32_t nc, const uint32_t* at,
                       uint32_t nl, uint32_t* out, hipStream_t s);
void launch_leaf_compact(DevCells c, uint32_t nc, const uint32_t* off, uint32_t* leaf_ids,
                         uint32_t* leaf_of_slot, hipStream_t s);

// ------------------------------------------------------------ upward.hip
bool launch_p2m(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                cplx* moments, hipStream_t s);
bool launch_m2m(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* moments,
                hipStream_t s);
void launch_pows(int p, uint32_t nc, const cplx* moments, const double* fact, double* pows,
                 hipStream_t s);
void launch_fprepass(uint32_t n, uint32_t nc, DevCells c, DevBodies b,
                     const uint64_t* keys_sorted, int vlev, double* f, hipStream_t s);
void launch_fmin(const uint32_t* cells, uint32_t cnt, DevCells c, const double* f, double* fmin,
                 hipStream_t s);

// -------------------------------------------------------------- walk.hip
struct MacParams {
  int p;
  int use_err;
  double eps;
  double theta2;
  const double* pows;
  const double* binom;
  const double* fmin;
  double soft2;     // softening^2
  double soft_mac;  // softening guard tau (0 = off)
};
size_t walk_rec_bytes();
// Device state of the breadth-first walk (walk.hip). Round r reads the
// frontier size fsize[r & 1] and writes fsize[(r + 1) & 1]; m2l_base /
// p2p_base are the list lengths at the start of the round being emitted.
// status bits 1 / 2 / 4 (frontier / M2L / P2P buffer too small for round
// `round`, need_* the sizes it needs) halt the walk at that round: the round
// and every later one leave all outputs and the state untouched, so the walk
// resumes at `round` once the buffers have grown.
struct WalkState {
  uint32_t fsize[2];
  uint32_t n_m2l, n_p2p;
  uint32_t m2l_base, p2p_base;
  uint32_t round;   // next round to run
  uint32_t rounds;  // rounds done with a non-empty frontier
  uint32_t status;
  uint32_t need_f, need_m2l, need_p2p;
};
constexpr uint32_t WALK_G = 2048;  // blocks per walk launch (blk holds 3 x WALK_G words)
void launch_front_init(const uint32_t* cells, uint32_t cnt, void* front, WalkState* st,
                       hipStream_t s);
// Walk records, and {cx, cy, cz, mass} of every cell as one 32-byte record.
void launch_pack_walk(uint32_t nc, DevCells c, void* rec, double4* cpm, hipStream_t s);
// Round r of the walk: decide, scan + state update, emit (fin / fout: the
// frontier buffers of round r; code holds at least the frontier capacity).
void launch_walk_round(const void* fin, void* fout, uint32_t r, WalkState* st, uint32_t cap_f,
                       uint2* m2l, uint32_t cap_m2l, uint2* p2p, uint32_t cap_p2p, uint16_t* code,
                       uint32_t* blk, const void* rec, const double* mass, MacParams mac,
                       hipStream_t s);
void launch_csr_ptr(const uint64_t* keys, uint32_t nnz, int sbits, uint32_t ntarget,
                    uint32_t* ptr, hipStream_t s);

// --------------------------------------------------------------- m2l.hip
// Chunks of at most `ch` sources per target (the M2L work unit).
size_t m2l_max_chunks(uint32_t nnz, uint32_t nc, uint32_t ch);
size_t m2l_part_bytes(int p, size_t maxchunks);
size_t m2l_chunk_bytes();
void m2l_build_chunks(const uint32_t* tptr, uint32_t nc, uint32_t ch, uint32_t* nch,
                      uint32_t* choff, uint32_t* scan_tmp, void* chunks, hipStream_t s,
                      uint32_t** total_dev);
bool launch_m2l_kernel(int p, const void* chunks, uint32_t nchunks, const uint64_t* skeys,
                       int sbits, DevCells c, const cplx* moments, double soft2, cplx* part,
                       hipStream_t s);
// Mixed form: the same M2L on radius-scaled FP32 operands (m2l_kernel.h);
// lam [nc] doubles and m32 [nc] x m2l32_moment_bytes(p) are scratch.
// prep_below: FP32 moments are made from the FP64 ones for the cells above
// this level only (the cells below hold them from the FP32 upward pass);
// 255 = every cell.
bool launch_m2l32(int p, uint32_t nc, const void* chunks, uint32_t nchunks, const uint64_t* skeys,
                  int sbits, DevCells c, const BBox* bb, const uint64_t* keys_sorted, int vlev,
                  const cplx* moments, double soft2, double* lam, void* m32, void* part,
                  hipStream_t s, int prep_below = 255);
// Per-cell scale of the mixed form: radius, floored at a quarter of the octant edge.
void launch_m2l_scale(uint32_t nc, DevCells c, const BBox* bb, const uint64_t* keys_sorted,
                      int vlev, double* lam, hipStream_t s);
size_t m2l32_moment_bytes(int p);
bool launch_m2l_reduce32(int p, uint32_t nc, const uint32_t* choff, const void* part,
                         const double* lam, cplx* locals, hipStream_t s);
bool launch_m2l_reduce(int p, uint32_t nc, const uint32_t* choff, const cplx* part, cplx* locals,
                       hipStream_t s);

// ----------------------------------------------------------- upward32.hip
// Mixed form's upward pass below the FP64 top levels: FP32 moments scaled by
// lam^-n (m >= 0 half, m32 [nc] x hnpos(p) complex floats); and their FP64
// view (sel >= 0: the cells at level sel and the leaves above it; sel < 0:
// every cell).
bool launch_p2m32(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                  const double* lam, void* m32, hipStream_t s);
bool launch_m2m32(int p, const uint32_t* cells, uint32_t cnt, DevCells c, const double* lam,
                  void* m32, hipStream_t s);
bool launch_m32_to_f64(int p, uint32_t nc, DevCells c, int sel, const double* lam,
                       const void* m32, cplx* out, hipStream_t s);
// Levels below a system root whose M2M stays FP64 in the mixed form.
constexpr int FP64_TOP = 3;

// ------------------------------------------- generic_up.hip / generic_far.hip
constexpr int TEMPLATED_MAX_P = 10;  // above this the runtime-p kernels run
void launch_p2m_gen(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                    cplx* moments, hipStream_t s);
void launch_m2m_gen(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* moments,
                    hipStream_t s);
void rot_tables_host(int p, std::vector<double>& sw, std::vector<double>& swt,
                     std::vector<cplx>& din, std::vector<cplx>& dout, std::vector<cplx>& fin,
                     std::vector<cplx>& fout, std::vector<double>& fact);
size_t rot_tables_struct_bytes();
void make_rot_tables(const double* sw, const double* swt, const cplx* din, const cplx* dout,
                     const cplx* fin, const cplx* fout, const double* fact, void* out);
void launch_m2l_gen(int p, uint32_t nc, const uint32_t* tptr, const uint64_t* skeys, int sbits,
                    DevCells c, const cplx* moments, double soft2, const void* rot_tables,
                    cplx* locals, hipStream_t s);
void launch_l2l_gen(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* locals,
                    hipStream_t s);
void launch_l2p_gen(int p, const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                    const cplx* locals, const double* lap, double4* far, hipStream_t s);

// ---------------------------------------------------------- downward.hip
bool launch_l2l(int p, const uint32_t* cells, uint32_t cnt, DevCells c, cplx* locals,
                hipStream_t s);
// loc32 non-null: the mixed form's scaled FP32 locals (with their scales lam).
bool launch_l2p(int p, uint32_t n, const uint32_t* leaf_of_slot, DevCells c, DevBodies b,
                const cplx* locals, const double* lap, double4* far, hipStream_t s,
                const double* lam = nullptr, const void* loc32 = nullptr);
// Mixed form: locals scaled by lam^n in FP32 (m >= 0 half, loc32 [nc] x
// hnpos(p) complex floats): one level's cells from their own FP64 locals and
// their parents' scaled locals; and the FP64 view of them.
bool launch_l2l32(int p, const uint32_t* cells, uint32_t cnt, DevCells c, const double* lam,
                  const cplx* own, void* loc32, hipStream_t s);
bool launch_loc32_export(int p, uint32_t nc, const double* lam, const void* loc32, cplx* out,
                         hipStream_t s);

// --------------------------------------------------------------- soft.hip
// Trace terms of the softened far field (lap may be null: eps = 0).
void launch_trace_leaf(const uint32_t* leaf_ids, uint32_t nleaves, DevCells c, DevBodies b,
                       double* q, hipStream_t s);
// The children's dipoles come from mom (FP64), or, for children deeper than
// level m32_below, from the mixed form's scaled FP32 moments m32 (with lam)
// when m32 is non-null.
void launch_trace_up(const uint32_t* cells, uint32_t cnt, DevCells c, int p, const cplx* mom,
                     double* q, hipStream_t s, const void* m32 = nullptr,
                     const double* lam = nullptr, int m32_below = 255);
void launch_m2l_trace(uint32_t nc, const uint32_t* tptr, const uint64_t* skeys, int sbits,
                      const double4* cpm, const double* q, double soft2, int p, cplx* locals,
                      double* lap, hipStream_t s);
// Target-side trace polynomial of the softened far field, summed over a
// leaf and all its ancestors (each cell's L, G, H about its own centre; body
// at (bx, by, bz)): pot += |d|^2 L/6 + |d|^2 (d.G)/10 + |d|^2 (d.H.d)/28
// - |d|^4 tr(H)/280 and acc += its gradient, d = body - centre.
constexpr int SOFT_LAP_WORDS = 10;  // L, G (3), H (xx, yy, zz, xy, xz, yz)
__device__ __forceinline__ void l2p_trace_walk(const double* __restrict__ lap,
                                               const int32_t* __restrict__ parent,
                                               const double* __restrict__ cx,
                                               const double* __restrict__ cy,
                                               const double* __restrict__ cz, uint32_t leaf,
                                               double bx, double by, double bz, double4& o) {
  // the polynomial's constant factors as reciprocals: a multiply each
  // instead of an FP64 division sequence per term, body and ancestor
  constexpr double C3 = 1.0 / 3.0, C6 = 1.0 / 6.0, C10 = 1.0 / 10.0, C14 = 1.0 / 14.0,
                   C28 = 1.0 / 28.0, C70 = 1.0 / 70.0, C280 = 1.0 / 280.0;
  for (int32_t A = (int32_t)leaf; A >= 0; A = parent[A]) {
    const double* t = lap + (size_t)SOFT_LAP_WORDS * A;
    const double dx = bx - cx[A], dy = by - cy[A], dz = bz - cz[A];
    const double L = t[0];
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double dg = dx * t[1] + dy * t[2] + dz * t[3];
    const double hx = t[4] * dx + t[7] * dy + t[8] * dz;
    const double hy = t[7] * dx + t[5] * dy + t[9] * dz;
    const double hz = t[8] * dx + t[9] * dy + t[6] * dz;
    const double dhd = dx * hx + dy * hy + dz * hz;
    const double tr = t[4] + t[5] + t[6];
    o.w += d2 * L * C6 + d2 * dg * C10 + d2 * dhd * C28 - d2 * d2 * tr * C280;
    o.x += dx * L * C3 + (2.0 * dx * dg + d2 * t[1]) * C10 + (dx * dhd + d2 * hx) * C14 -
           d2 * dx * tr * C70;
    o.y += dy * L * C3 + (2.0 * dy * dg + d2 * t[2]) * C10 + (dy * dhd + d2 * hy) * C14 -
           d2 * dy * tr * C70;
    o.z += dz * L * C3 + (2.0 * dz * dg + d2 * t[3]) * C10 + (dz * dhd + d2 * hz) * C14 -
           d2 * dz * tr * C70;
  }
}

// -------------------------------------------------------------- near.hip
// P2P work items over the leaves (see near.hip): items / reds / partial
// live in the context; nitems >= nleaves, nsplit = leaves cut into several
// items.
struct P2PPlan {
  const uint4* items;
  uint32_t nitems;
  const uint4* reds;
  uint32_t nsplit;
  double4* partial;
};
// item_work: a leaf whose targets x sources exceed it is cut into items.
void p2p_plan_count(const uint32_t* leaf_ids, uint32_t nleaves, const uint32_t* pptr,
                    const uint64_t* skeys, int sbits, const uint32_t* ccount, uint64_t item_work,
                    uint32_t* nitem, uint32_t* npart, uint32_t* nsplit, hipStream_t s);
void p2p_plan_emit(const uint32_t* leaf_ids, uint32_t nleaves, const uint32_t* pptr,
                   const uint64_t* skeys, int sbits, const uint32_t* ccount, uint64_t item_work,
                   const uint32_t* ioff, const uint32_t* poff, const uint32_t* roff, uint4* items,
                   uint4* reds, hipStream_t s);
// skeys: the target-sorted directed P2P entries (source in the low sbits)
void launch_p2p(const P2PPlan& pl, DevCells c, const double4* body, const uint64_t* skeys,
                int sbits, double soft2, const double4* far, const uint32_t* order,
                double* accel, double* pot, hipStream_t s);
void launch_p2p_f32(const P2PPlan& pl, DevCells c, const double4* body, const uint64_t* skeys,
                    int sbits, double soft2, const double4* far, const uint32_t* order,
                    double* accel, double* pot, hipStream_t s);
// Leaves cut into several items: their partial sums and the far field, added
// in item order (compensated in the mixed form) into accel / pot.
void launch_p2p_reduce(const P2PPlan& pl, DevCells c, const double4* far, const uint32_t* order,
                       double* accel, double* pot, bool compensated, hipStream_t s);
// Directed body-body interactions of a P2P pair list (self pairs c(c-1));
// out[0] = total, out needs 1 + p2p_count_scratch(np) words.
size_t p2p_count_scratch(uint32_t np);
void launch_p2p_count(const uint2* pairs, uint32_t np, const uint32_t* count,
                      unsigned long long* out, hipStream_t s);

// --------------------------------------------------------- integrate.hip
// st[0] = dt of the current step, st[1] = time advanced (device scalars).
size_t dt_scratch_words();
void launch_step_dt(const double* acc, size_t n, double dt_req, double eta, double eps_len,
                    double* part, double* st, hipStream_t s);
void launch_kick_drift(size_t n, double* pos, double* vel, const double* acc, const double* st,
                       hipStream_t s);
void launch_kick(size_t n, double* vel, const double* acc, const double* st, hipStream_t s);
void launch_cap_accel(size_t n, double* acc, double cap, hipStream_t s);
void launch_fill_u64(void* p, size_t words, uint64_t v, hipStream_t s);
void launch_drift_check(uint32_t n, const uint32_t* leaf_of_slot, DevCells c, DevBodies b,
                        const BBox* bb, const uint64_t* keys_sorted, int vlev, double tol,
                        uint32_t* flag, hipStream_t s);

// ------------------------------------------------------------ direct.hip
// FP64 direct sum at nt sampled targets: each target's sources are cut into
// nsplit contiguous ranges (one block each, partial sums in part[nt][nsplit][5]),
// then added in range order into acc [nt][3], pot [nt], f [nt] (pot, f may be
// null). direct_split picks nsplit for (n, nt) within the allocation-free
// entry point's scratch, so both entry points sum in the same order.
uint32_t direct_split(size_t n, size_t nt);
void launch_direct(const double* pos, const double* mass, size_t n, double soft2,
                   const uint32_t* ids, uint32_t nt, uint32_t nsplit, double* part, double* acc,
                   double* pot, double* f, hipStream_t s);

}  // namespace fmmh
