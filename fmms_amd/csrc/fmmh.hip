// fmmh.hip — host orchestration and the C ABI (include/fmmh.h).
//
// One context owns a HIP stream (default priority) and a device arena. A
// build runs the Morton tree on the device (tree.hip); an evaluation runs
// the upward pass, the optional f pre-pass, the dual walk, the target-grouped
// M2L / L2L / L2P and the fused P2P + scatter. Sizes that decide launch
// shapes (cells per level, walk frontier, list lengths) are read back once
// per level / round; everything else stays on the device.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

using namespace fmmh;

namespace {

enum Stage {
  ST_TREE = 0,
  ST_UPWARD,
  ST_PREPASS,
  ST_WALK,
  ST_LISTS,
  ST_M2L,
  ST_M2L_REDUCE,
  ST_L2L,
  ST_L2P,
  ST_P2P,  // the pair kernel alone
  ST_P2P_REDUCE,
  ST_COUNT
};
const char* kStageNames[ST_COUNT] = {"tree", "upward", "prepass",    "walk", "lists", "m2l",
                                     "m2l_reduce", "l2l", "l2p", "p2p", "p2p_reduce"};

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

struct HipError {
  std::string msg;
};

#define HCHECK(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw HipError{std::string(#call) + ": " + hipGetErrorString(e_)};         \
  } while (0)

struct Budget {
  size_t limit;
};

}  // namespace

struct fmmh_ctx {
  fmmh_params prm;
  fmmh_tuning tun = {32, (uint64_t)1 << 18};
  hipStream_t s = nullptr;
  std::string err;
  size_t used = 0;
  std::vector<Buf*> bufs;
  // ensemble layout: nsys systems with body ranges sys_off[s]..sys_off[s+1];
  // vlev virtual tree levels select the system (0 for one system)
  uint32_t nsys = 1;
  int vlev = 0;
  std::vector<uint64_t> sys_off;
  Buf d_sysoff;
  // state
  bool built = false, evaluated = false;
  size_t n = 0;
  uint32_t nc = 0, nleaves = 0, max_level = 0;
  int sbits = 1;
  std::vector<uint32_t> lvl_off;
  uint64_t* keys_sorted = nullptr;
  uint32_t* order = nullptr;
  uint32_t n_m2l = 0, n_p2p = 0, walk_rounds = 0;
  uint32_t walk_batch = 8;  // walk rounds launched per host check
  uint64_t p2p_interactions = 0;  // directed body-body pairs of the lists (on request)
  bool p2p_counted = false;
  // timing
  hipEvent_t ev[ST_COUNT + 1] = {};
  bool ev_init = false;
  double stage_ms[ST_COUNT] = {};
  // buffers
  Buf in_pos, in_mass, out_acc, out_pot;
  Buf keys, keys_alt, vals, vals_alt;
  Buf bx, by, bz, bm, bxyzm;
  Buf bpack;  // (x, y, z, m) in id order, written with the keys for the gather
  Buf bbox_partial, bbox, flags;
  Buf chain, top, cstart;
  Buf bfs_first, bfs_count, bfs_parent, bfs_oct, bfs_level, bfs_dfs, nchild, child_off;
  Buf c_first, c_count, c_children, c_parent, c_leaf, c_level, c_cx, c_cy, c_cz, c_mass, c_rho,
      c_rope, c_last;
  Buf lvl_cells, leaf_flags, leaf_off, leaf_ids, leaf_of_slot;
  Buf int_pos, int_cells, lvl_dev;  // internal cells per level (BFS order)
  Buf split_bnd;                    // octant boundaries of a level's cells (9 words each)
  Buf leaf_bc, lbc_keys, lbc_keys_alt;  // leaves ordered by body count
  Buf cpm;                              // {cx, cy, cz, mass} per cell
  std::vector<uint32_t> int_off;   // level L's internal cells: [int_off[L], int_off[L+1])
  Buf scan_tmp, sort_tmp;
  Buf moments, locals, pows, fvals, fmin, fact, binom;
  Buf front_a, front_b, m2l, p2p, ctr;
  Buf mkeys, mkeys_alt, tptr, pkeys, pkeys_alt, pptr, list_tot;
  Buf nch, choff, chunks, part;
  Buf far, p2p_work, walkrec;
  Buf walk_code, walk_blk;  // the walk's per-pair decisions, per-block counts
  Buf pl_n, pl_items, pl_reds, pl_partial;  // P2P work plan
  Buf m2l_lam, m2l_m32;                     // mixed form: FP32 M2L scales, scaled moments
  Buf loc32;                                // mixed form: FP32 locals scaled by lam^n
  uint32_t p_nitems = 0, p_nsplit = 0;
  Buf rt_sw, rt_swt, rt_din, rt_dout, rt_fin, rt_fout, rt_fact;
  Buf ctrace, clap;  // softened far field: q = sum m |x - z|^2; L, G, H per cell
  int rot_p = -1;
  std::vector<char> rot_struct;
  Buf direct_ids, direct_out, direct_part;
  Buf pinned;

  fmmh_ctx() {
    Buf* all[] = {&in_pos, &in_mass, &out_acc, &out_pot, &keys, &keys_alt, &vals, &vals_alt,
                  &bx, &by, &bz, &bm, &bxyzm, &bpack, &bbox_partial, &bbox, &flags, &chain, &top, &cstart,
                  &bfs_first, &bfs_count, &bfs_parent, &bfs_oct, &bfs_level, &bfs_dfs, &nchild,
                  &child_off, &c_first, &c_count, &c_children, &c_parent, &c_leaf, &c_level,
                  &c_cx, &c_cy, &c_cz, &c_mass, &c_rho, &c_rope, &c_last, &lvl_cells,
                  &leaf_flags, &leaf_off, &leaf_ids, &leaf_of_slot, &int_pos, &int_cells, &lvl_dev, &split_bnd, &leaf_bc, &lbc_keys, &lbc_keys_alt, &cpm,
                  &scan_tmp, &sort_tmp,
                  &moments, &locals, &pows, &fvals, &fmin, &fact, &binom, &front_a, &front_b,
                  &m2l, &p2p, &ctr, &mkeys, &mkeys_alt, &tptr, &pkeys, &pkeys_alt, &pptr, &list_tot,
                  &nch, &choff, &chunks, &part, &far, &p2p_work, &walkrec, &walk_code, &walk_blk, &pl_n, &pl_items,
                  &pl_reds, &pl_partial, &m2l_lam, &m2l_m32, &loc32, &rt_sw, &rt_swt, &rt_din, &rt_dout, &rt_fin, &rt_fout,
                  &rt_fact, &direct_ids, &direct_out, &direct_part, &lf_part, &lf_state, &lf_flag, &ctrace, &clap,
                  &d_sysoff};
    bufs.assign(std::begin(all), std::end(all));
  }

  ~fmmh_ctx() {
    drop_graph();
    for (Buf* b : bufs)
      if (b->p) (void)hipFree(b->p);
    if (pinned.p) (void)hipHostFree(pinned.p);
    if (ev_init)
      for (auto& e : ev) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
  }

  // Grow a device buffer to at least `bytes` (content kept when keep=true).
  template <class T = void>
  T* ensure(Buf& b, size_t bytes, bool keep = false) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return (T*)b.p;
    size_t want = std::max(bytes, b.cap + b.cap / 2);
    size_t limit = (size_t)prm.hbm_budget_bytes;
    if (limit && used - b.cap + want > limit) {
      want = bytes;
      if (used - b.cap + want > limit) throw Budget{limit};
    }
    void* np = nullptr;
    HCHECK(hipMalloc(&np, want));
    if (keep && b.p) HCHECK(hipMemcpyAsync(np, b.p, b.cap, hipMemcpyDeviceToDevice, s));
    if (b.p) {
      HCHECK(hipStreamSynchronize(s));
      HCHECK(hipFree(b.p));
    }
    used = used - b.cap + want;
    b.p = np;
    b.cap = want;
    return (T*)np;
  }

  void* host_stage(size_t bytes) {
    if (pinned.cap < bytes) {
      if (pinned.p) HCHECK(hipHostFree(pinned.p));
      pinned.p = nullptr;
      size_t want = std::max<size_t>(bytes, 4096);
      HCHECK(hipHostMalloc(&pinned.p, want, hipHostMallocDefault));
      pinned.cap = want;
    }
    return pinned.p;
  }

  template <class T>
  T read1(const T* dev) {
    T* h = (T*)host_stage(sizeof(T));
    HCHECK(hipMemcpyAsync(h, dev, sizeof(T), hipMemcpyDeviceToHost, s));
    HCHECK(hipStreamSynchronize(s));
    return *h;
  }

  void mark(int st) {
    if (!(prm.flags & FMMH_FLAG_TIMING) || capturing) return;
    if (!ev_init) {
      for (auto& e : ev) HCHECK(hipEventCreate(&e));
      ev_init = true;
    }
    HCHECK(hipEventRecord(ev[st], s));
  }

  void finish_timing(int first, int last) {
    if (!(prm.flags & FMMH_FLAG_TIMING) || !ev_init) return;
    HCHECK(hipStreamSynchronize(s));
    for (int st = first; st < last; st++) {
      float ms = 0.f;
      HCHECK(hipEventElapsedTime(&ms, ev[st], ev[st + 1]));
      stage_ms[st] = ms;
    }
  }

  DevCells cells() {
    return DevCells{(uint32_t*)c_first.p, (uint32_t*)c_count.p, (int32_t*)c_children.p,
                    (int32_t*)c_parent.p, (uint8_t*)c_leaf.p,  (uint8_t*)c_level.p,
                    (double*)c_cx.p,      (double*)c_cy.p,     (double*)c_cz.p,
                    (double*)c_mass.p,    (double*)c_rho.p,    (int32_t*)c_rope.p,
                    (int32_t*)c_last.p};
  }
  DevBodies bodies() {
    return DevBodies{(double*)bx.p, (double*)by.p, (double*)bz.p, (double*)bm.p, (double4*)bxyzm.p};
  }
  DevBfs bfs() {
    return DevBfs{(uint32_t*)bfs_first.p, (uint32_t*)bfs_count.p, (int32_t*)bfs_parent.p,
                  (uint8_t*)bfs_oct.p,    (uint8_t*)bfs_level.p,  (uint32_t*)bfs_dfs.p};
  }

  void ensure_bfs(size_t cap, bool keep) {
    ensure(bfs_first, cap * 4, keep);
    ensure(bfs_count, cap * 4, keep);
    ensure(bfs_parent, cap * 4, keep);
    ensure(bfs_oct, cap, keep);
    ensure(bfs_level, cap, keep);
    ensure(bfs_dfs, cap * 4, false);
  }

  void ensure_scan(size_t n_elems) {
    ensure(scan_tmp, scan_tmp_words(n_elems) * 4 + 64);
  }

  // ------------------------------------------------------------------ build
  void build(size_t n_, const double* pos, const double* mass) {
    n = n_;
    built = evaluated = lists_ok = false;
    tree_evals = 0;
    drop_graph();
    const int p = prm.p;
    (void)p;
    ensure(keys, n * 8);
    ensure(keys_alt, n * 8);
    ensure(vals, n * 4);
    ensure(vals_alt, n * 4);
    ensure(bx, n * 8);
    ensure(by, n * 8);
    ensure(bz, n * 8);
    ensure(bm, n * 8);
    ensure(bxyzm, n * 32);
    ensure(bpack, n * 32);
    ensure(bbox_partial, 6 * 1024 * 8);
    ensure(bbox, nsys * sizeof(BBox));
    ensure(flags, 64);
    ensure(chain, n * 4);
    ensure(top, n);
    ensure(cstart, n * 4);
    ensure(sort_tmp, radix_tmp_words(n) * 4 + 64);
    ensure_scan(n);
    mark(ST_TREE);
    HCHECK(hipMemsetAsync(flags.p, 0, 64, s));
    uint64_t* off = nullptr;
    if (nsys == 1) {
      launch_bbox(pos, n, (double*)bbox_partial.p, (BBox*)bbox.p, s);
      launch_keys(pos, mass, n, (const BBox*)bbox.p, (uint64_t*)keys.p, (uint32_t*)vals.p,
                  (uint32_t*)flags.p, (double4*)bpack.p, s);
    } else {
      off = ensure<uint64_t>(d_sysoff, (nsys + 1) * 8);
      uint64_t* h = (uint64_t*)host_stage((nsys + 1) * 8);
      std::memcpy(h, sys_off.data(), (nsys + 1) * 8);
      HCHECK(hipMemcpyAsync(off, h, (nsys + 1) * 8, hipMemcpyHostToDevice, s));
      launch_bbox_sys(pos, off, nsys, (BBox*)bbox.p, s);
      launch_keys_sys(pos, mass, n, off, nsys, vlev, (const BBox*)bbox.p, (uint64_t*)keys.p,
                      (uint32_t*)vals.p, (uint32_t*)flags.p, (double4*)bpack.p, s);
    }
    int flip = radix_sort_u64((uint64_t*)keys.p, (uint32_t*)vals.p, (uint64_t*)keys_alt.p,
                              (uint32_t*)vals_alt.p, n, 0, 63, (uint32_t*)sort_tmp.p, s);
    keys_sorted = (uint64_t*)(flip ? keys_alt.p : keys.p);
    order = (uint32_t*)(flip ? vals_alt.p : vals.p);
    if (nsys > 1) {
      // bodies now in standalone-key order: take the system prefix, then a
      // stable sort on the prefix bits gives (system, standalone key) order
      launch_keys_prefix(keys_sorted, order, n, off, nsys, vlev, s);
      uint64_t* k2 = (uint64_t*)(flip ? keys.p : keys_alt.p);
      uint32_t* v2 = (uint32_t*)(flip ? vals.p : vals_alt.p);
      const int f2 = radix_sort_u64(keys_sorted, order, k2, v2, n, 63 - 3 * vlev, 63,
                                    (uint32_t*)sort_tmp.p, s);
      if (f2) {
        keys_sorted = k2;
        order = v2;
      }
    }
    launch_gather_packed((const double4*)bpack.p, order, n, bodies(), s);
    // level-by-level octant splits
    HCHECK(hipMemsetAsync(chain.p, 0, n * 4, s));
    HCHECK(hipMemsetAsync(top.p, 0, n, s));
    size_t cap = std::max<size_t>(1024, 4 * (n / std::max(1, prm.n_leaf)) + 64);
    ensure_bfs(cap, false);
    ensure(nchild, cap * 4);
    ensure(child_off, cap * 4);
    launch_tree_init(bfs(), n, (uint32_t*)chain.p, (uint8_t*)top.p, s);
    if (read1((const uint32_t*)flags.p) != 0) throw int(FMMH_E_NONFINITE);
    lvl_off.assign({0u, 1u});
    for (int L = 0;; L++) {
      uint32_t s0 = lvl_off[L], cnt = lvl_off[L + 1] - s0;
      ensure(nchild, (size_t)cnt * 4);
      ensure(child_off, (size_t)cnt * 4);
      ensure(split_bnd, (size_t)cnt * 36);
      ensure_scan(cnt);
      launch_split_count(bfs(), s0, cnt, keys_sorted, (uint32_t)prm.n_leaf, vlev,
                         (uint32_t*)nchild.p, (uint32_t*)split_bnd.p, s);
      uint32_t* tot = nullptr;
      scan_exclusive_u32((uint32_t*)nchild.p, (uint32_t*)child_off.p, cnt, (uint32_t*)scan_tmp.p, s,
                         &tot);
      uint32_t total = read1(tot);
      if (total == 0) break;
      size_t need = (size_t)lvl_off[L + 1] + total;
      if (need > bfs_first.cap / 4) ensure_bfs(need + need / 2, true);
      launch_split_emit(bfs(), s0, cnt, lvl_off[L + 1], (uint32_t)prm.n_leaf, vlev,
                        (uint32_t*)split_bnd.p, (uint32_t*)child_off.p, (uint32_t*)chain.p,
                        (uint8_t*)top.p, s);
      lvl_off.push_back((uint32_t)need);
    }
    nc = lvl_off.back();
    max_level = (uint32_t)lvl_off.size() - 2;
    scan_exclusive_u32((uint32_t*)chain.p, (uint32_t*)cstart.p, n, (uint32_t*)scan_tmp.p, s, nullptr);
    launch_dfs_ids(bfs(), nc, (uint32_t*)cstart.p, (uint8_t*)top.p, s);
    ensure(c_first, nc * 4);
    ensure(c_count, nc * 4);
    ensure(c_children, (size_t)nc * 32);
    ensure(c_parent, nc * 4);
    ensure(c_leaf, nc);
    ensure(c_level, nc);
    ensure(c_cx, nc * 8);
    ensure(c_cy, nc * 8);
    ensure(c_cz, nc * 8);
    ensure(c_mass, nc * 8);
    ensure(c_rho, nc * 8);
    ensure(c_rope, nc * 4);
    ensure(c_last, nc * 4);
    ensure(lvl_cells, nc * 4);
    ensure(leaf_flags, nc * 4);
    ensure(leaf_off, nc * 4);
    ensure(leaf_ids, nc * 4);
    ensure(leaf_of_slot, n * 4);
    ensure_scan(nc);
    HCHECK(hipMemsetAsync(c_children.p, 0xFF, (size_t)nc * 32, s));
    launch_scatter_cells(bfs(), nc, (uint32_t)prm.n_leaf, vlev, cells(), (uint32_t*)lvl_cells.p, s);
    launch_leaf_flags(cells(), nc, (uint32_t*)leaf_flags.p, s);
    uint32_t* tot = nullptr;
    scan_exclusive_u32((uint32_t*)leaf_flags.p, (uint32_t*)leaf_off.p, nc, (uint32_t*)scan_tmp.p, s,
                       &tot);
    launch_leaf_compact(cells(), nc, (uint32_t*)leaf_off.p, (uint32_t*)leaf_ids.p,
                        (uint32_t*)leaf_of_slot.p, s);
    nleaves = read1(tot);
    internal_lists();
    ensure(leaf_bc, (size_t)nleaves * 4 + 4);
    ensure(lbc_keys, (size_t)nleaves * 8 + 8);
    ensure(lbc_keys_alt, (size_t)nleaves * 8 + 8);
    ensure(sort_tmp, radix_tmp_words(nleaves) * 4 + 64);
    launch_leaves_by_count((uint32_t*)leaf_ids.p, nleaves, cells(), (uint64_t*)lbc_keys.p,
                           (uint64_t*)lbc_keys_alt.p, (uint32_t*)sort_tmp.p, (uint32_t*)leaf_bc.p,
                           s);
    // geometry: every leaf at once (from its bodies), then internal cells
    // bottom-up level by level (from their children)
    launch_geometry_leaves((uint32_t*)leaf_ids.p, nleaves, cells(), bodies(), prm.center_mode, s);
    for (int L = (int)max_level - 1; L >= 0; L--)
      launch_geometry((uint32_t*)lvl_cells.p + lvl_off[L], lvl_off[L + 1] - lvl_off[L], cells(),
                      bodies(), prm.center_mode, s, 2);
    for (int L = 0; L <= (int)max_level; L++)
      launch_ropes((uint32_t*)lvl_cells.p + lvl_off[L], lvl_off[L + 1] - lvl_off[L], cells(), s);
    ensure(walkrec, (size_t)nc * walk_rec_bytes());
    ensure(cpm, (size_t)nc * sizeof(double4));
    launch_pack_walk(nc, cells(), walkrec.p, (double4*)cpm.p, s);
    sbits = 1;
    while ((1ull << sbits) <= nc) sbits++;
    built = true;
  }

  // Internal cells of each level, for the per-level upward kernels.
  void internal_lists() {
    ensure(int_pos, (size_t)nc * 4);
    ensure(int_cells, (size_t)nc * 4);
    launch_internal_flags((uint32_t*)lvl_cells.p, nc, cells(), (uint32_t*)leaf_flags.p, s);
    uint32_t* itot = nullptr;
    scan_exclusive_u32((uint32_t*)leaf_flags.p, (uint32_t*)int_pos.p, nc, (uint32_t*)scan_tmp.p, s,
                       &itot);
    launch_internal_compact((uint32_t*)lvl_cells.p, nc, cells(), (uint32_t*)int_pos.p,
                            (uint32_t*)int_cells.p, s);
    const uint32_t nl = (uint32_t)lvl_off.size();
    ensure(lvl_dev, (size_t)nl * 8);
    uint32_t* h = (uint32_t*)host_stage((size_t)nl * 4);
    std::memcpy(h, lvl_off.data(), (size_t)nl * 4);
    HCHECK(hipMemcpyAsync(lvl_dev.p, h, (size_t)nl * 4, hipMemcpyHostToDevice, s));
    launch_level_pick((uint32_t*)int_pos.p, itot, nc, (uint32_t*)lvl_dev.p, nl,
                      (uint32_t*)lvl_dev.p + nl, s);
    HCHECK(hipMemcpyAsync(h, (uint32_t*)lvl_dev.p + nl, (size_t)nl * 4, hipMemcpyDeviceToHost, s));
    HCHECK(hipStreamSynchronize(s));
    int_off.assign(h, h + nl);
  }

  // --------------------------------------------------------------- evaluate
  // Lists of the last full evaluation, kept for frozen evaluations until the
  // next build (directed M2L / P2P keys sorted by target, CSR, M2L chunks).
  bool lists_ok = false;
  const uint64_t* msorted = nullptr;
  const uint64_t* psorted = nullptr;
  uint32_t nchunks = 0;
  uint32_t tree_evals = 0;  // evaluations run on the current tree and lists
  bool capturing = false;

  // Gather (optional) + P2M + M2M into `moments` (already sized); in the
  // mixed form with the theta MAC, FP32 scaled moments below the FP64 top
  // levels go straight into m2l_m32.
  void upward(const double* pos, const double* mass, bool regather) {
    const int p = prm.p;
    const size_t D = (size_t)hdim(p);
    DevCells c = cells();
    DevBodies b = bodies();
    if (regather) launch_gather(pos, mass, order, n, b, s);
    mark(ST_UPWARD);
    cplx* mom = (cplx*)moments.p;
    if (up32()) {
      double* lam = ensure<double>(m2l_lam, (size_t)nc * sizeof(double));
      void* m32 = ensure(m2l_m32, (size_t)nc * m2l32_moment_bytes(p));
      const int ltop = vlev + FP64_TOP;
      launch_m2l_scale(nc, c, (const BBox*)bbox.p, keys_sorted, vlev, lam, s);
      launch_p2m32(p, (uint32_t*)leaf_bc.p, nleaves, c, b, lam, m32, s);
      for (int L = (int)max_level - 1; L >= ltop; L--)
        launch_m2m32(p, (uint32_t*)int_cells.p + int_off[L], int_off[L + 1] - int_off[L], c, lam,
                     m32, s);
      launch_m32_to_f64(p, nc, c, ltop, lam, m32, mom, s);  // children of the top levels
      for (int L = std::min(ltop, (int)max_level) - 1; L >= 0; L--)
        launch_m2m(p, (uint32_t*)int_cells.p + int_off[L], int_off[L + 1] - int_off[L], c, mom, s);
    } else if (p <= TEMPLATED_MAX_P) {
      launch_p2m(p, (uint32_t*)leaf_bc.p, nleaves, c, b, mom, s);
      for (int L = (int)max_level - 1; L >= 0; L--)
        launch_m2m(p, (uint32_t*)int_cells.p + int_off[L], int_off[L + 1] - int_off[L], c, mom, s);
    } else {
      launch_fill_u64(mom, nc * D * 2, 0, s);  // kernel fill: safe under graph capture
      launch_p2m_gen(p, (uint32_t*)leaf_ids.p, nleaves, c, b, mom, s);
      for (int L = (int)max_level - 1; L >= 0; L--)
        launch_m2m_gen(p, (uint32_t*)lvl_cells.p + lvl_off[L], lvl_off[L + 1] - lvl_off[L], c, mom,
                       s);
    }
    if (prm.softening > 0.0) {  // trace moments of the softened far field
      double* q = (double*)ctrace.p;
      launch_trace_leaf((uint32_t*)leaf_ids.p, nleaves, c, b, q, s);
      for (int L = (int)max_level - 1; L >= 0; L--)
        launch_trace_up((uint32_t*)int_cells.p + int_off[L], int_off[L + 1] - int_off[L], c, p, mom,
                        q, s, up32() ? m2l_m32.p : nullptr, (const double*)m2l_lam.p,
                        vlev + FP64_TOP);
    }
  }

  // Multipole powers, f pre-pass and fmin (error-controlled MAC only).
  void mac_inputs() {
    const int p = prm.p;
    DevCells c = cells();
    DevBodies b = bodies();
    std::vector<double> f(2 * p + 1, 1.0);
    for (size_t i = 1; i < f.size(); i++) f[i] = f[i - 1] * (double)i;  // moments.rs:11-17
    std::vector<double> bn(p + 1, 1.0);
    for (int k = 0; k <= p; k++) bn[k] = f[p] / (f[k] * f[p - k]);  // moments.rs:20-27
    double* hf = (double*)host_stage((f.size() + bn.size()) * 8);
    std::memcpy(hf, f.data(), f.size() * 8);
    std::memcpy(hf + f.size(), bn.data(), bn.size() * 8);
    ensure(fact, f.size() * 8);
    ensure(binom, bn.size() * 8);
    HCHECK(hipMemcpyAsync(fact.p, hf, f.size() * 8, hipMemcpyHostToDevice, s));
    HCHECK(hipMemcpyAsync(binom.p, hf + f.size(), bn.size() * 8, hipMemcpyHostToDevice, s));
    HCHECK(hipStreamSynchronize(s));
    ensure(fmin, nc * 8);
    ensure(pows, (size_t)nc * (p + 1) * 8);
    ensure(fvals, n * 8);
    launch_pows(p, nc, (const cplx*)moments.p, (double*)fact.p, (double*)pows.p, s);
    launch_fprepass((uint32_t)n, nc, c, b, keys_sorted, vlev, (double*)fvals.p, s);
    for (int L = (int)max_level; L >= 0; L--)
      launch_fmin((uint32_t*)lvl_cells.p + lvl_off[L], lvl_off[L + 1] - lvl_off[L], c,
                  (double*)fvals.p, (double*)fmin.p, s);
  }

  // Directed lists grouped by target, P2P work count, M2L chunks.
  void build_lists() {
    const int p = prm.p;
    ensure(mkeys, (size_t)2 * n_m2l * 8);
    ensure(mkeys_alt, (size_t)2 * n_m2l * 8);
    ensure(tptr, ((size_t)nc + 1) * 4);
    ensure(sort_tmp, radix_tmp_words((size_t)2 * n_m2l) * 4 + 64);
    // The walk writes its pairs in one fixed order, so a stable sort on the
    // target bits alone groups the directed entries deterministically (each
    // target's sources in walk order).
    int fl = radix_sort_directed((const uint2*)m2l.p, n_m2l, sbits, false, (uint64_t*)mkeys.p,
                                 (uint64_t*)mkeys_alt.p, (uint32_t*)sort_tmp.p, s);
    msorted = (const uint64_t*)(fl ? mkeys_alt.p : mkeys.p);
    launch_csr_ptr(msorted, 2 * n_m2l, sbits, nc, (uint32_t*)tptr.p, s);
    ensure(pkeys, (size_t)2 * n_p2p * 8);
    ensure(pkeys_alt, (size_t)2 * n_p2p * 8);
    ensure(pptr, ((size_t)nc + 1) * 4);
    ensure(sort_tmp, radix_tmp_words((size_t)2 * n_p2p) * 4 + 64);
    // a self pair's second entry is a sentinel that sorts past every real
    // entry: the row pointers stop before them, and ptr[nc] counts the rest
    fl = radix_sort_directed((const uint2*)p2p.p, n_p2p, sbits, true, (uint64_t*)pkeys.p,
                             (uint64_t*)pkeys_alt.p, (uint32_t*)sort_tmp.p, s);
    psorted = (const uint64_t*)(fl ? pkeys_alt.p : pkeys.p);
    launch_csr_ptr(psorted, 2 * n_p2p, sbits, nc, (uint32_t*)pptr.p, s);
    p2p_counted = false;
    // the P2P plan's three totals and the M2L chunk count, copied into one
    // small buffer on the device and read back together
    ensure(list_tot, 16);
    uint32_t* dtot = (uint32_t*)list_tot.p;
    HCHECK(hipMemsetAsync(dtot, 0, 16, s));
    plan_p2p_count(dtot);
    if (p <= TEMPLATED_MAX_P) {
      const uint32_t ch = (uint32_t)tun.m2l_chunk;
      size_t maxck = m2l_max_chunks(2 * n_m2l, nc, ch);
      ensure(nch, ((size_t)nc + 1) * 4);
      ensure(choff, ((size_t)nc + 1) * 4);
      ensure(chunks, maxck * m2l_chunk_bytes());
      ensure_scan(nc + 1);
      uint32_t* tot = nullptr;
      m2l_build_chunks((uint32_t*)tptr.p, nc, ch, (uint32_t*)nch.p, (uint32_t*)choff.p,
                       (uint32_t*)scan_tmp.p, chunks.p, s, &tot);
      // choff[nc] = total chunks (the reduce reads choff[t+1])
      HCHECK(hipMemcpyAsync((uint32_t*)choff.p + nc, tot, 4, hipMemcpyDeviceToDevice, s));
      HCHECK(hipMemcpyAsync(dtot + 3, tot, 4, hipMemcpyDeviceToDevice, s));
    }
    const uint4 t = read1((const uint4*)dtot);
    p_nitems = t.x;
    p_nsplit = t.z;
    plan_p2p_emit(t.y);
    if (p <= TEMPLATED_MAX_P) {
      nchunks = t.w;
      ensure(part, m2l_part_bytes(p, std::max<size_t>(nchunks, 1)));
      if (prm.precision == FMMH_PREC_FP32BULK) {
        ensure(m2l_lam, (size_t)nc * sizeof(double));
        ensure(m2l_m32, (size_t)nc * m2l32_moment_bytes(p));
        ensure(loc32, (size_t)nc * m2l32_moment_bytes(p));
      }
    } else if (p >= 20) {
      rot_tables();
    }
    ensure(far, n * sizeof(double4));
  }

  // P2P work items (near.hip): per leaf the item count, partial-sum slots and
  // split flag, scanned in place into offsets (the three totals to
  // tot[0..2]), then, once the totals are on the host, the items themselves.
  void plan_p2p_count(uint32_t* tot) {
    p_nitems = p_nsplit = 0;
    if (nleaves == 0) return;
    const size_t w = (size_t)nleaves + 1;
    ensure(pl_n, 3 * w * 4);
    uint32_t* nitem = (uint32_t*)pl_n.p;
    p2p_plan_count((const uint32_t*)leaf_ids.p, nleaves, (const uint32_t*)pptr.p, psorted, sbits,
                   (const uint32_t*)c_count.p, tun.p2p_item_work, nitem, nitem + w, nitem + 2 * w,
                   s);
    ensure_scan(w);
    for (int k = 0; k < 3; k++) {
      uint32_t* col = nitem + k * w;
      uint32_t* t = nullptr;
      scan_exclusive_u32(col, col, w, (uint32_t*)scan_tmp.p, s, &t);
      HCHECK(hipMemcpyAsync(tot + k, t, 4, hipMemcpyDeviceToDevice, s));
    }
  }

  void plan_p2p_emit(uint32_t nslots) {
    if (nleaves == 0) return;
    const size_t w = (size_t)nleaves + 1;
    uint32_t* nitem = (uint32_t*)pl_n.p;
    ensure(pl_items, (size_t)std::max<uint32_t>(p_nitems, 1) * sizeof(uint4));
    ensure(pl_reds, (size_t)std::max<uint32_t>(p_nsplit, 1) * sizeof(uint4));
    ensure(pl_partial, (size_t)std::max<uint32_t>(nslots, 1) * sizeof(double4));
    p2p_plan_emit((const uint32_t*)leaf_ids.p, nleaves, (const uint32_t*)pptr.p, psorted, sbits,
                  (const uint32_t*)c_count.p, tun.p2p_item_work, nitem, nitem + w, nitem + 2 * w,
                  (uint4*)pl_items.p, (uint4*)pl_reds.p, s);
  }

  // The mixed form's FP32 M2L, L2L and L2P (templated orders only).
  bool mixed32() const { return prm.precision == FMMH_PREC_FP32BULK && prm.p <= TEMPLATED_MAX_P; }
  // ... and its FP32 P2M / M2M below the top levels (theta MAC: the error MAC
  // reads the FP64 moments of every cell).
  bool up32() const { return mixed32() && !std::isfinite(prm.eps_mac); }

  // M2L (+ reduce), L2L, L2P and the fused P2P + scatter into acc / pot.
  // Launches only: safe to capture into a graph once the lists exist.
  void far_near(double* acc, double* pot) {
    const int p = prm.p;
    DevCells c = cells();
    DevBodies b = bodies();
    const cplx* mom = (const cplx*)moments.p;
    cplx* loc = (cplx*)locals.p;
    const double soft2 = prm.softening * prm.softening;
    const bool gen = p > TEMPLATED_MAX_P;
    mark(ST_M2L);
    if (!gen && prm.precision == FMMH_PREC_FP32BULK) {
      launch_m2l32(p, nc, chunks.p, nchunks, msorted, sbits, c, (const BBox*)bbox.p, keys_sorted,
                   vlev, mom, soft2, (double*)m2l_lam.p, m2l_m32.p, part.p, s,
                   up32() ? vlev + FP64_TOP : 255);
      mark(ST_M2L_REDUCE);
      launch_m2l_reduce32(p, nc, (uint32_t*)choff.p, part.p, (const double*)m2l_lam.p, loc, s);
    } else if (!gen) {
      launch_m2l_kernel(p, chunks.p, nchunks, msorted, sbits, c, mom, soft2, (cplx*)part.p, s);
      mark(ST_M2L_REDUCE);
      launch_m2l_reduce(p, nc, (uint32_t*)choff.p, (cplx*)part.p, loc, s);
    } else {
      launch_m2l_gen(p, nc, (uint32_t*)tptr.p, msorted, sbits, c, mom, soft2, rot_struct.data(),
                     loc, s);
      mark(ST_M2L_REDUCE);
    }
    double* lap = nullptr;
    if (soft2 > 0.0) {
      lap = (double*)clap.p;
      launch_m2l_trace(nc, (const uint32_t*)tptr.p, msorted, sbits, (const double4*)cpm.p,
                       (const double*)ctrace.p, soft2, p, loc, lap, s);
    }
    mark(ST_L2L);
    const bool l32 = mixed32();
    for (int L = l32 ? 0 : 1; L <= (int)max_level; L++) {
      const uint32_t* lc = (uint32_t*)lvl_cells.p + lvl_off[L];
      const uint32_t cnt = lvl_off[L + 1] - lvl_off[L];
      if (l32)
        launch_l2l32(p, lc, cnt, c, (const double*)m2l_lam.p, loc, loc32.p, s);
      else if (!gen)
        launch_l2l(p, lc, cnt, c, loc, s);
      else
        launch_l2l_gen(p, lc, cnt, c, loc, s);
    }
    mark(ST_L2P);
    if (l32)
      launch_l2p(p, (uint32_t)n, (uint32_t*)leaf_of_slot.p, c, b, loc, lap, (double4*)far.p, s,
                 (const double*)m2l_lam.p, loc32.p);
    else if (!gen)
      launch_l2p(p, (uint32_t)n, (uint32_t*)leaf_of_slot.p, c, b, loc, lap, (double4*)far.p, s);
    else
      launch_l2p_gen(p, (uint32_t*)leaf_ids.p, nleaves, c, b, loc, lap, (double4*)far.p, s);
    const P2PPlan pl{(const uint4*)pl_items.p, p_nitems, (const uint4*)pl_reds.p, p_nsplit,
                     (double4*)pl_partial.p};
    const bool f32 = prm.precision == FMMH_PREC_FP32BULK;
    mark(ST_P2P);
    if (f32)
      launch_p2p_f32(pl, c, b.xyzm, psorted, sbits, soft2, (const double4*)far.p, order, acc, pot,
                     s);
    else
      launch_p2p(pl, c, b.xyzm, psorted, sbits, soft2, (const double4*)far.p, order, acc, pot, s);
    mark(ST_P2P_REDUCE);
    launch_p2p_reduce(pl, c, (const double4*)far.p, order, acc, pot, f32, s);
    mark(ST_COUNT);
  }

  void evaluate(const double* pos, const double* mass, double* acc, double* pot, bool regather) {
    const size_t D = (size_t)hdim(prm.p);
    rot_struct.resize(std::max(rot_struct.size(), rot_tables_struct_bytes()));
    lists_ok = false;
    drop_graph();
    ensure(moments, nc * D * sizeof(cplx));
    ensure(locals, nc * D * sizeof(cplx));
    if (prm.softening > 0.0) {
      ensure(ctrace, nc * 8);
      ensure(clap, nc * SOFT_LAP_WORDS * sizeof(double));
    }
    upward(pos, mass, regather);
    mark(ST_PREPASS);
    const bool use_err = std::isfinite(prm.eps_mac);
    if (use_err) mac_inputs();
    else ensure(fmin, nc * 8);
    mark(ST_WALK);
    walk(use_err);
    mark(ST_LISTS);
    build_lists();
    far_near(acc, pot);
    HCHECK(hipGetLastError());
    HCHECK(hipStreamSynchronize(s));  // every call returns with its results in place
    evaluated = true;
    lists_ok = true;
    tree_evals = 1;
  }

  // Temporal reuse (gpufmm.rs:2710-2741 frozen_step): bodies re-gathered
  // into the built tree's slots, expansions recomputed about the frozen cell
  // centres, interactions from the last full evaluation's lists.
  void evaluate_frozen(const double* pos, const double* mass, double* acc, double* pot) {
    upward(pos, mass, true);
    mark(ST_PREPASS);
    mark(ST_WALK);
    mark(ST_LISTS);
    far_near(acc, pot);
    tree_evals++;
  }

  // ----------------------------------------------------------- leapfrog
  Buf lf_part, lf_state, lf_flag;
  uint64_t lf_stats[4] = {0, 0, 0, 0};  // steps, rebuilds, frozen steps, drift-forced rebuilds
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  struct GraphKey {
    const void *pos, *vel, *mass, *acc, *pot;
    double dt, eta, eps_len, cap, tol;
    bool operator==(const GraphKey& o) const {
      return pos == o.pos && vel == o.vel && mass == o.mass && acc == o.acc && pot == o.pot &&
             dt == o.dt && eta == o.eta && eps_len == o.eps_len && cap == o.cap && tol == o.tol;
    }
  } gkey{};

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  // One step on the current tree and lists: dt, kick + drift, frozen
  // evaluation, optional cap, kick (sim.rs:548-572).
  void frozen_step(double* pos, double* vel, const double* mass, double* acc, double* pot,
                   const fmmh_leapfrog_opts& o, double* part, double* st) {
    launch_step_dt(acc, n, o.dt, o.dt_eta, o.dt_eps_len, part, st, s);
    launch_kick_drift(n, pos, vel, acc, st, s);
    evaluate_frozen(pos, mass, acc, pot);
    if (o.drift_tol > 0.0)
      launch_drift_check((uint32_t)n, (const uint32_t*)leaf_of_slot.p, cells(), bodies(),
                         (const BBox*)bbox.p, keys_sorted, vlev, o.drift_tol,
                         (uint32_t*)lf_flag.p, s);
    if (o.accel_cap > 0.0) launch_cap_accel(n, acc, o.accel_cap, s);
    launch_kick(n, vel, acc, st, s);
  }

  void capture_frozen_step(double* pos, double* vel, const double* mass, double* acc, double* pot,
                           const fmmh_leapfrog_opts& o, double* part, double* st) {
    drop_graph();
    const uint32_t age = tree_evals;  // capture records the step, it does not run it
    capturing = true;
    HCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    frozen_step(pos, vel, mass, acc, pot, o, part, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    capturing = false;
    tree_evals = age;
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      throw HipError{std::string("hipStreamEndCapture: ") + hipGetErrorString(e)};
    }
    graph = g;
    HCHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    gkey = GraphKey{pos, vel, mass, acc, pot, o.dt, o.dt_eta, o.dt_eps_len, o.accel_cap,
                    o.drift_tol};
  }

  double leapfrog(size_t n_, double* pos, double* vel, const double* mass, double* acc,
                  double* pot, const fmmh_leapfrog_opts& o) {
    double* part = ensure<double>(lf_part, dt_scratch_words() * 8);
    double* st = ensure<double>(lf_state, 64);
    uint32_t* flag = ensure<uint32_t>(lf_flag, 64);
    launch_fill_u64(st, 2, 0, s);
    bool drifted = false;
    auto full = [&] {
      build(n_, pos, mass);
      evaluate(pos, mass, acc, pot, false);
      if (o.accel_cap > 0.0) launch_cap_accel(n, acc, o.accel_cap, s);
      launch_fill_u64(flag, 1, 0, s);
      drifted = false;
      lf_stats[1]++;
    };
    if (!(o.flags & FMMH_LF_ACC_VALID)) {
      full();
    } else if (!lists_ok || n != n_) {
      // accelerations given, no current tree: the first step rebuilds
      n = n_;
      tree_evals = (uint32_t)o.refresh_every;
    }
    const bool use_graph = (o.flags & FMMH_LF_GRAPH) != 0;
    const GraphKey key{pos, vel, mass, acc, pot, o.dt, o.dt_eta, o.dt_eps_len, o.accel_cap,
                       o.drift_tol};
    for (int k = 0; k < o.steps; k++) {
      lf_stats[0]++;
      if (drifted || tree_evals >= (uint32_t)o.refresh_every) {
        if (drifted && tree_evals < (uint32_t)o.refresh_every) lf_stats[3]++;
        launch_step_dt(acc, n, o.dt, o.dt_eta, o.dt_eps_len, part, st, s);
        launch_kick_drift(n, pos, vel, acc, st, s);
        full();
        launch_kick(n, vel, acc, st, s);
        continue;
      }
      if (use_graph) {
        if (!gexec || !(gkey == key)) capture_frozen_step(pos, vel, mass, acc, pot, o, part, st);
        HCHECK(hipGraphLaunch(gexec, s));
        tree_evals++;
      } else {
        frozen_step(pos, vel, mass, acc, pot, o, part, st);
      }
      lf_stats[2]++;
      // a body left its frozen leaf: the next step rebuilds the tree and lists
      if (o.drift_tol > 0.0 && k + 1 < o.steps) drifted = read1(flag) != 0;
    }
    if (o.drift_tol > 0.0) drifted = read1(flag) != 0;
    if (drifted) tree_evals = (uint32_t)o.refresh_every;  // carried into the next call
    double* h = (double*)host_stage(16);
    HCHECK(hipMemcpyAsync(h, st, 16, hipMemcpyDeviceToHost, s));
    HCHECK(hipStreamSynchronize(s));
    HCHECK(hipGetLastError());
    return h[1];
  }

  // Rotation tables of the current p on the device (built once per p).
  void rot_tables() {
    const int p = prm.p;
    if (rot_p == p) return;
    std::vector<double> sw, swt, fact;
    std::vector<cplx> din, dout, fin, fout;
    rot_tables_host(p, sw, swt, din, dout, fin, fout, fact);
    auto up = [&](Buf& b, const void* src, size_t bytes) {
      ensure(b, bytes);
      HCHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    up(rt_sw, sw.data(), sw.size() * 8);
    up(rt_swt, swt.data(), swt.size() * 8);
    up(rt_din, din.data(), din.size() * sizeof(cplx));
    up(rt_dout, dout.data(), dout.size() * sizeof(cplx));
    up(rt_fin, fin.data(), fin.size() * sizeof(cplx));
    up(rt_fout, fout.data(), fout.size() * sizeof(cplx));
    up(rt_fact, fact.data(), fact.size() * 8);
    rot_struct.resize(rot_tables_struct_bytes());
    make_rot_tables((double*)rt_sw.p, (double*)rt_swt.p, (cplx*)rt_din.p, (cplx*)rt_dout.p,
                    (cplx*)rt_fin.p, (cplx*)rt_fout.p, (double*)rt_fact.p, rot_struct.data());
    rot_p = p;
  }

  // The dual walk, breadth-first (walk.hip): per round the decisions and
  // their per-block counts, then the outputs in frontier order; one readback
  // per round gives the next frontier's size. An output buffer that ran out
  // of room grows and the round's outputs are written again.
  void walk(bool use_err) {
    const size_t cap0 = (size_t)4 * nc + 1024;
    if (front_a.cap < cap0 * 8) {
      ensure(front_a, cap0 * 8);
      ensure(front_b, cap0 * 8);
    }
    ensure(m2l, (size_t)(32 * (size_t)nc + 1024) * 8);
    ensure(p2p, (size_t)(16 * (size_t)nc + 1024) * 8);
    ensure(ctr, 64);
    MacParams mp;
    mp.p = prm.p;
    mp.use_err = use_err ? 1 : 0;
    mp.eps = prm.eps_mac;
    mp.theta2 = prm.theta_max * prm.theta_max;
    mp.pows = (const double*)pows.p;
    mp.binom = (const double*)binom.p;
    mp.fmin = (const double*)fmin.p;
    mp.soft2 = prm.softening * prm.softening;
    mp.soft_mac = prm.soft_mac;
    // the walk starts from the self pair of every (system) root: level vlev
    const uint32_t nf0 = lvl_off[vlev + 1] - lvl_off[vlev];
    WalkState* st = (WalkState*)ctr.p;
    HCHECK(hipMemsetAsync(st, 0, sizeof(WalkState), s));
    launch_front_init((const uint32_t*)lvl_cells.p + lvl_off[vlev], nf0, front_a.p, st, s);
    ensure(walk_blk, (size_t)3 * WALK_G * 4);
    Buf* fr[2] = {&front_a, &front_b};
    auto cap = [](const Buf& b) { return (uint32_t)std::min<size_t>(b.cap / 8, 0xFFFFFFFFu); };
    // Rounds are launched in batches (as many as the last walk took, plus
    // one) and the host reads the walk state once per batch: the walk is
    // done when the frontier of the next round is empty; a halted round
    // (a buffer too small) resumes after the buffers have grown.
    uint32_t r = 0;
    for (;;) {
      ensure(walk_code, (size_t)std::max(cap(front_a), cap(front_b)) * 2);
      const uint32_t batch = std::max<uint32_t>(4u, walk_batch);
      for (uint32_t k = 0; k < batch; k++, r++)
        launch_walk_round(fr[r & 1u]->p, fr[(r + 1u) & 1u]->p, r, st, cap(*fr[(r + 1u) & 1u]),
                          (uint2*)m2l.p, cap(m2l), (uint2*)p2p.p, cap(p2p),
                          (uint16_t*)walk_code.p, (uint32_t*)walk_blk.p, walkrec.p,
                          (const double*)c_mass.p, mp, s);
      const WalkState h = read1((const WalkState*)st);
      r = h.round;
      if (h.status != 0u) {
        // round r needs more room: its input frontier is kept, the rest grows
        if (h.status & 1u) {
          ensure(*fr[(r + 1u) & 1u], (size_t)h.need_f * 8 * 2);
          ensure(*fr[r & 1u], fr[(r + 1u) & 1u]->cap, true);
        }
        if (h.status & 2u) ensure(m2l, (size_t)h.need_m2l * 8 * 2, true);
        if (h.status & 4u) ensure(p2p, (size_t)h.need_p2p * 8 * 2, true);
        HCHECK(hipMemsetAsync(&st->status, 0, sizeof(uint32_t), s));
        continue;
      }
      if (h.fsize[r & 1u] == 0u) {
        n_m2l = h.n_m2l;
        n_p2p = h.n_p2p;
        walk_rounds = h.rounds;
        walk_batch = h.rounds + 1u;
        break;
      }
    }
  }
};

namespace {

int fail(fmmh_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int check_params(const fmmh_params* p, std::string* why) {
  if (p->p < 1 || p->p > FMMH_MAX_P) {
    *why = "p must be in [1, " + std::to_string(FMMH_MAX_P) + "]";
    return FMMH_E_PARAM;
  }
  if (p->n_leaf < 1) {
    *why = "n_leaf must be >= 1";
    return FMMH_E_PARAM;
  }
  if (!(p->theta_max > 0.0 && p->theta_max <= 1.0)) {
    *why = "theta_max must be in (0, 1]";
    return FMMH_E_PARAM;
  }
  if (!(p->eps_mac > 0.0)) {
    *why = "eps_mac must be > 0 (INFINITY disables the error test)";
    return FMMH_E_PARAM;
  }
  if (!(p->softening >= 0.0) || !std::isfinite(p->softening)) {
    *why = "softening must be finite and >= 0";
    return FMMH_E_PARAM;
  }
  if (!(p->soft_mac >= 0.0) || !std::isfinite(p->soft_mac)) {
    *why = "soft_mac must be finite and >= 0";
    return FMMH_E_PARAM;
  }
  if (p->center_mode != FMMH_CENTER_COM && p->center_mode != FMMH_CENTER_SPHERE) {
    *why = "center_mode must be FMMH_CENTER_COM or FMMH_CENTER_SPHERE";
    return FMMH_E_PARAM;
  }
  if (p->precision != FMMH_PREC_FP64 && p->precision != FMMH_PREC_FP32BULK) {
    *why = "precision must be FMMH_PREC_FP64 or FMMH_PREC_FP32BULK";
    return FMMH_E_PARAM;
  }
  return FMMH_OK;
}

template <class F>
int guarded(fmmh_ctx* c, F fn) {
  if (!c) return FMMH_E_STATE;
  try {
    fn();
    return FMMH_OK;
  } catch (const HipError& e) {
    return fail(c, FMMH_E_HIP, e.msg);
  } catch (const Budget& b) {
    return fail(c, FMMH_E_OVER_BUDGET,
                "device memory budget of " + std::to_string(b.limit) + " bytes exceeded");
  } catch (int code) {
    return fail(c, code, code == FMMH_E_NONFINITE ? "non-finite position or mass" : "error");
  }
}

// models.rs:13-31 SplitMix64
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) / (double)(1ull << 53); }
};

// sin and cos of one angle through glibc's sincos: the route LLVM (the
// reference's Rust) and GCC take for a sin/cos pair of the same argument;
// separate sin() and cos() calls can differ from it in the last bit.
inline void sin_cos(double a, double* s, double* c) { ::sincos(a, s, c); }

}  // namespace

extern "C" {

void fmmh_default_params(fmmh_params* out) {
  out->p = 8;
  out->eps_mac = 1e-5;
  out->theta_max = 1.0;
  out->n_leaf = 16;
  out->softening = 0.0;
  out->center_mode = FMMH_CENTER_COM;
  out->precision = FMMH_PREC_FP64;
  out->flags = 0;
  out->hbm_budget_bytes = 0.0;
  out->soft_mac = 0.0;
}

int fmmh_create(const fmmh_params* params, fmmh_ctx** out) {
  if (!params || !out) return FMMH_E_PARAM;
  *out = nullptr;
  std::string why;
  int rc = check_params(params, &why);
  if (rc != FMMH_OK) return rc;
  fmmh_ctx* c = new fmmh_ctx();
  c->prm = *params;
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FMMH_E_HIP;
  }
  *out = c;
  return FMMH_OK;
}

void fmmh_destroy(fmmh_ctx* ctx) { delete ctx; }

const char* fmmh_last_error(const fmmh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fmmh_set_systems(fmmh_ctx* ctx, size_t nsys, const uint64_t* offsets) {
  if (!ctx) return FMMH_E_STATE;
  if (nsys == 0 || nsys > FMMH_MAX_SYSTEMS)
    return fail(ctx, FMMH_E_PARAM, "nsys must be in [1, " + std::to_string(FMMH_MAX_SYSTEMS) + "]");
  if (!offsets || offsets[0] != 0) return fail(ctx, FMMH_E_PARAM, "offsets[0] must be 0");
  for (size_t i = 0; i < nsys; i++)
    if (offsets[i + 1] < offsets[i]) return fail(ctx, FMMH_E_PARAM, "offsets must not decrease");
  if (offsets[nsys] >= 0xFFFFFFFFull) return fail(ctx, FMMH_E_PARAM, "n must be < 2^32");
  ctx->nsys = (uint32_t)nsys;
  ctx->vlev = 0;
  while ((1ull << (3 * ctx->vlev)) < nsys) ctx->vlev++;
  ctx->sys_off.assign(offsets, offsets + nsys + 1);
  ctx->built = ctx->evaluated = ctx->lists_ok = false;
  ctx->drop_graph();
  return FMMH_OK;
}

static int check_layout(fmmh_ctx* ctx, size_t n) {
  if (ctx->nsys > 1 && n != ctx->sys_off[ctx->nsys])
    return fail(ctx, FMMH_E_PARAM, "n differs from the ensemble's offsets[nsys]");
  return FMMH_OK;
}

int fmmh_build_device(fmmh_ctx* ctx, size_t n, const double* d_pos, const double* d_mass) {
  if (!ctx) return FMMH_E_STATE;
  if (n == 0) return fail(ctx, FMMH_E_EMPTY, "empty input (n == 0)");
  if (int rc = check_layout(ctx, n)) return rc;
  if (n >= 0xFFFFFFFFull) return fail(ctx, FMMH_E_PARAM, "n must be < 2^32");
  return guarded(ctx, [&] {
    ctx->build(n, d_pos, d_mass);
    ctx->finish_timing(ST_TREE, ST_TREE);
  });
}

int fmmh_evaluate_device(fmmh_ctx* ctx, size_t n, const double* d_pos, const double* d_mass,
                         double* d_accel, double* d_pot) {
  if (!ctx) return FMMH_E_STATE;
  if (!ctx->built) return fail(ctx, FMMH_E_STATE, "evaluate before build");
  if (n != ctx->n) return fail(ctx, FMMH_E_STATE, "n differs from the built tree");
  return guarded(ctx, [&] {
    ctx->evaluate(d_pos, d_mass, d_accel, d_pot, true);
    ctx->stage_ms[ST_TREE] = 0.0;
    ctx->finish_timing(ST_UPWARD, ST_COUNT);
  });
}

int fmmh_accel_device(fmmh_ctx* ctx, size_t n, const double* d_pos, const double* d_mass,
                      double* d_accel, double* d_pot) {
  if (!ctx) return FMMH_E_STATE;
  if (n == 0) return fail(ctx, FMMH_E_EMPTY, "empty input (n == 0)");
  if (int rc = check_layout(ctx, n)) return rc;
  if (n >= 0xFFFFFFFFull) return fail(ctx, FMMH_E_PARAM, "n must be < 2^32");
  return guarded(ctx, [&] {
    ctx->build(n, d_pos, d_mass);
    ctx->evaluate(d_pos, d_mass, d_accel, d_pot, false);
    ctx->finish_timing(ST_TREE, ST_COUNT);
  });
}

int fmmh_evaluate_frozen_device(fmmh_ctx* ctx, size_t n, const double* d_pos,
                                const double* d_mass, double* d_accel, double* d_pot) {
  if (!ctx) return FMMH_E_STATE;
  if (!ctx->lists_ok) return fail(ctx, FMMH_E_STATE, "frozen evaluation needs a full evaluation");
  if (n != ctx->n) return fail(ctx, FMMH_E_STATE, "n differs from the built tree");
  return guarded(ctx, [&] {
    ctx->evaluate_frozen(d_pos, d_mass, d_accel, d_pot);
    HCHECK(hipGetLastError());
    ctx->stage_ms[ST_TREE] = 0.0;
    ctx->finish_timing(ST_UPWARD, ST_COUNT);
    if (!(ctx->prm.flags & FMMH_FLAG_TIMING)) HCHECK(hipStreamSynchronize(ctx->s));
  });
}

void fmmh_leapfrog_default(fmmh_leapfrog_opts* o) {
  o->dt = 1e-4;
  o->steps = 1;
  o->refresh_every = 1;
  o->dt_eta = 0.0;
  o->dt_eps_len = 0.02;
  o->accel_cap = 0.0;
  o->drift_tol = 0.25;
  o->flags = 0;
}

int fmmh_leapfrog_stats(const fmmh_ctx* ctx, uint64_t* out4) {
  if (!ctx || !out4) return FMMH_E_STATE;
  for (int i = 0; i < 4; i++) out4[i] = ctx->lf_stats[i];
  return FMMH_OK;
}

int fmmh_leapfrog_device(fmmh_ctx* ctx, size_t n, double* d_pos, double* d_vel,
                         const double* d_mass, double* d_accel, double* d_pot,
                         const fmmh_leapfrog_opts* o, double* time_advanced) {
  if (!ctx) return FMMH_E_STATE;
  if (!o || !d_pos || !d_vel || !d_mass || !d_accel || !d_pot)
    return fail(ctx, FMMH_E_PARAM, "null argument");
  if (n == 0) return fail(ctx, FMMH_E_EMPTY, "empty input (n == 0)");
  if (n >= 0xFFFFFFFFull) return fail(ctx, FMMH_E_PARAM, "n must be < 2^32");
  if (int rc = check_layout(ctx, n)) return rc;
  if (!(o->dt > 0.0) || !std::isfinite(o->dt) || o->steps < 0 || o->refresh_every < 1 ||
      !(o->dt_eta >= 0.0) || !(o->dt_eps_len > 0.0) || !(o->accel_cap >= 0.0) ||
      !(o->drift_tol >= 0.0) ||
      (o->flags & ~(FMMH_LF_ACC_VALID | FMMH_LF_GRAPH)))
    return fail(ctx, FMMH_E_PARAM,
                "leapfrog options: dt > 0, steps >= 0, refresh_every >= 1, dt_eta >= 0, "
                "dt_eps_len > 0, accel_cap >= 0, drift_tol >= 0, known flags only");
  return guarded(ctx, [&] {
    const double t = ctx->leapfrog(n, d_pos, d_vel, d_mass, d_accel, d_pot, *o);
    if (time_advanced) *time_advanced = t;
  });
}

static int upload(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass) {
  return guarded(ctx, [&] {
    ctx->ensure(ctx->in_pos, n * 24);
    ctx->ensure(ctx->in_mass, n * 8);
    HCHECK(hipMemcpyAsync(ctx->in_pos.p, pos, n * 24, hipMemcpyHostToDevice, ctx->s));
    HCHECK(hipMemcpyAsync(ctx->in_mass.p, mass, n * 8, hipMemcpyHostToDevice, ctx->s));
  });
}

static int download(fmmh_ctx* ctx, size_t n, double* accel, double* pot) {
  return guarded(ctx, [&] {
    HCHECK(hipMemcpyAsync(accel, ctx->out_acc.p, n * 24, hipMemcpyDeviceToHost, ctx->s));
    if (pot) HCHECK(hipMemcpyAsync(pot, ctx->out_pot.p, n * 8, hipMemcpyDeviceToHost, ctx->s));
    HCHECK(hipStreamSynchronize(ctx->s));
  });
}

int fmmh_build(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass) {
  if (!ctx) return FMMH_E_STATE;
  if (n == 0) return fail(ctx, FMMH_E_EMPTY, "empty input (n == 0)");
  int rc = upload(ctx, n, pos, mass);
  if (rc) return rc;
  return fmmh_build_device(ctx, n, (const double*)ctx->in_pos.p, (const double*)ctx->in_mass.p);
}

static int outputs(fmmh_ctx* ctx, size_t n) {
  return guarded(ctx, [&] {
    ctx->ensure(ctx->out_acc, n * 24);
    ctx->ensure(ctx->out_pot, n * 8);
  });
}

int fmmh_evaluate(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass, double* accel,
                  double* pot) {
  if (!ctx) return FMMH_E_STATE;
  if (!ctx->built) return fail(ctx, FMMH_E_STATE, "evaluate before build");
  if (n != ctx->n) return fail(ctx, FMMH_E_STATE, "n differs from the built tree");
  int rc = upload(ctx, n, pos, mass);
  if (!rc) rc = outputs(ctx, n);
  if (!rc)
    rc = fmmh_evaluate_device(ctx, n, (const double*)ctx->in_pos.p, (const double*)ctx->in_mass.p,
                              (double*)ctx->out_acc.p, (double*)ctx->out_pot.p);
  if (!rc) rc = download(ctx, n, accel, pot);
  return rc;
}

int fmmh_accel(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass, double* accel,
               double* pot) {
  if (!ctx) return FMMH_E_STATE;
  if (n == 0) return fail(ctx, FMMH_E_EMPTY, "empty input (n == 0)");
  int rc = upload(ctx, n, pos, mass);
  if (!rc) rc = outputs(ctx, n);
  if (!rc)
    rc = fmmh_accel_device(ctx, n, (const double*)ctx->in_pos.p, (const double*)ctx->in_mass.p,
                           (double*)ctx->out_acc.p, (double*)ctx->out_pot.p);
  if (!rc) rc = download(ctx, n, accel, pot);
  return rc;
}

int fmmh_evaluate_frozen(fmmh_ctx* ctx, size_t n, const double* pos, const double* mass,
                         double* accel, double* pot) {
  if (!ctx) return FMMH_E_STATE;
  if (!ctx->lists_ok) return fail(ctx, FMMH_E_STATE, "frozen evaluation needs a full evaluation");
  if (n != ctx->n) return fail(ctx, FMMH_E_STATE, "n differs from the built tree");
  int rc = upload(ctx, n, pos, mass);
  if (!rc) rc = outputs(ctx, n);
  if (!rc)
    rc = fmmh_evaluate_frozen_device(ctx, n, (const double*)ctx->in_pos.p,
                                     (const double*)ctx->in_mass.p, (double*)ctx->out_acc.p,
                                     (double*)ctx->out_pot.p);
  if (!rc) rc = download(ctx, n, accel, pot);
  return rc;
}

void fmmh_default_tuning(fmmh_tuning* out) {
  out->m2l_chunk = 32;
  out->p2p_item_work = (uint64_t)1 << 18;
}

int fmmh_set_tuning(fmmh_ctx* ctx, const fmmh_tuning* t) {
  if (!ctx) return FMMH_E_STATE;
  if (!t || t->m2l_chunk < 1 || t->m2l_chunk > 1024 || t->p2p_item_work < 64 ||
      t->p2p_item_work > ((uint64_t)1 << 40))
    return fail(ctx, FMMH_E_PARAM, "tuning: m2l_chunk in [1, 1024], p2p_item_work in [64, 2^40]");
  ctx->tun = *t;
  ctx->lists_ok = false;  // the chunks and the P2P plan follow at the next full evaluation
  ctx->drop_graph();
  return FMMH_OK;
}

int fmmh_get_tuning(const fmmh_ctx* ctx, fmmh_tuning* out) {
  if (!ctx || !out) return FMMH_E_STATE;
  *out = ctx->tun;
  return FMMH_OK;
}

size_t fmmh_num_cells(const fmmh_ctx* ctx) { return ctx && ctx->built ? ctx->nc : 0; }

size_t fmmh_device_bytes(const fmmh_ctx* ctx) { return ctx ? ctx->used : 0; }

int fmmh_get_tree(fmmh_ctx* ctx, uint32_t* order, uint64_t* keys_sorted, uint32_t* first,
                  uint32_t* count, int32_t* children, uint8_t* leaf, uint32_t* level, double* cen,
                  double* mass, double* rho) {
  if (!ctx || !ctx->built) return fail(ctx, FMMH_E_STATE, "no tree");
  return guarded(ctx, [&] {
    const size_t n = ctx->n, nc = ctx->nc;
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
      if (dst) HCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->s));
    };
    d2h(order, ctx->order, n * 4);
    d2h(keys_sorted, ctx->keys_sorted, n * 8);
    d2h(first, ctx->c_first.p, nc * 4);
    d2h(count, ctx->c_count.p, nc * 4);
    d2h(children, ctx->c_children.p, nc * 32);
    d2h(leaf, ctx->c_leaf.p, nc);
    d2h(mass, ctx->c_mass.p, nc * 8);
    d2h(rho, ctx->c_rho.p, nc * 8);
    std::vector<uint8_t> lv(nc);
    std::vector<double> x(nc), y(nc), z(nc);
    if (level) d2h(lv.data(), ctx->c_level.p, nc);
    if (cen) {
      d2h(x.data(), ctx->c_cx.p, nc * 8);
      d2h(y.data(), ctx->c_cy.p, nc * 8);
      d2h(z.data(), ctx->c_cz.p, nc * 8);
    }
    HCHECK(hipStreamSynchronize(ctx->s));
    if (level)
      for (size_t i = 0; i < nc; i++) level[i] = lv[i];
    if (cen)
      for (size_t i = 0; i < nc; i++) {
        cen[3 * i] = x[i];
        cen[3 * i + 1] = y[i];
        cen[3 * i + 2] = z[i];
      }
  });
}

int fmmh_get_list_sizes(fmmh_ctx* ctx, size_t* n_m2l, size_t* n_p2p) {
  if (!ctx || !ctx->evaluated) return fail(ctx, FMMH_E_STATE, "no evaluation");
  if (n_m2l) *n_m2l = ctx->n_m2l;
  if (n_p2p) *n_p2p = ctx->n_p2p;
  return FMMH_OK;
}

int fmmh_get_lists(fmmh_ctx* ctx, uint32_t* m2l, uint32_t* p2p) {
  if (!ctx || !ctx->evaluated) return fail(ctx, FMMH_E_STATE, "no evaluation");
  return guarded(ctx, [&] {
    if (m2l && ctx->n_m2l)
      HCHECK(hipMemcpyAsync(m2l, ctx->m2l.p, (size_t)ctx->n_m2l * 8, hipMemcpyDeviceToHost, ctx->s));
    if (p2p && ctx->n_p2p)
      HCHECK(hipMemcpyAsync(p2p, ctx->p2p.p, (size_t)ctx->n_p2p * 8, hipMemcpyDeviceToHost, ctx->s));
    HCHECK(hipStreamSynchronize(ctx->s));
  });
}

static int get_dev(fmmh_ctx* ctx, void* out, const void* src, size_t bytes) {
  if (!ctx || !ctx->evaluated) return fail(ctx, FMMH_E_STATE, "no evaluation");
  return guarded(ctx, [&] {
    HCHECK(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, ctx->s));
    HCHECK(hipStreamSynchronize(ctx->s));
  });
}

int fmmh_get_moments(fmmh_ctx* ctx, double* out) {
  if (!ctx) return FMMH_E_STATE;
  if (ctx->evaluated && ctx->up32()) {  // FP64 view of the scaled FP32 moments
    int rc = guarded(ctx, [&] {
      const size_t bytes = (size_t)ctx->nc * hdim(ctx->prm.p) * 16;
      ctx->ensure(ctx->direct_out, bytes);
      launch_m32_to_f64(ctx->prm.p, ctx->nc, ctx->cells(), -1, (const double*)ctx->m2l_lam.p,
                        ctx->m2l_m32.p, (cplx*)ctx->direct_out.p, ctx->s);
    });
    if (rc) return rc;
    return get_dev(ctx, out, ctx->direct_out.p, (size_t)ctx->nc * hdim(ctx->prm.p) * 16);
  }
  return get_dev(ctx, out, ctx->moments.p, (size_t)ctx->nc * hdim(ctx->prm.p) * 16);
}
int fmmh_get_locals(fmmh_ctx* ctx, double* out) {
  if (!ctx) return FMMH_E_STATE;
  if (ctx->evaluated && ctx->mixed32()) {  // FP64 view of the scaled FP32 locals
    int rc = guarded(ctx, [&] {
      const size_t bytes = (size_t)ctx->nc * hdim(ctx->prm.p) * 16;
      ctx->ensure(ctx->direct_out, bytes);
      launch_loc32_export(ctx->prm.p, ctx->nc, (const double*)ctx->m2l_lam.p, ctx->loc32.p,
                          (cplx*)ctx->direct_out.p, ctx->s);
    });
    if (rc) return rc;
    return get_dev(ctx, out, ctx->direct_out.p, (size_t)ctx->nc * hdim(ctx->prm.p) * 16);
  }
  return get_dev(ctx, out, ctx->locals.p, (size_t)ctx->nc * hdim(ctx->prm.p) * 16);
}
int fmmh_get_pows(fmmh_ctx* ctx, double* out) {
  if (!ctx) return FMMH_E_STATE;
  if (!std::isfinite(ctx->prm.eps_mac)) return fail(ctx, FMMH_E_STATE, "theta-only MAC: no pows");
  return get_dev(ctx, out, ctx->pows.p, (size_t)ctx->nc * (ctx->prm.p + 1) * 8);
}
int fmmh_get_f(fmmh_ctx* ctx, double* f_slot, double* fmin) {
  if (!ctx) return FMMH_E_STATE;
  if (!std::isfinite(ctx->prm.eps_mac)) return fail(ctx, FMMH_E_STATE, "theta-only MAC: no f");
  int rc = f_slot ? get_dev(ctx, f_slot, ctx->fvals.p, ctx->n * 8) : FMMH_OK;
  if (!rc && fmin) rc = get_dev(ctx, fmin, ctx->fmin.p, (size_t)ctx->nc * 8);
  return rc;
}

int fmmh_stage_times(const fmmh_ctx* ctx, double* ms, int max) {
  if (!ctx) return 0;
  for (int i = 0; i < ST_COUNT && i < max; i++) ms[i] = ctx->stage_ms[i];
  return ST_COUNT;
}

const char* fmmh_stage_name(int stage) {
  return (stage >= 0 && stage < ST_COUNT) ? kStageNames[stage] : "";
}

int fmmh_work_counts(const fmmh_ctx* ctx, uint64_t* out8) {
  if (!ctx) return FMMH_E_STATE;
  if (ctx->lists_ok && !ctx->p2p_counted) {
    // counted on request, not in every evaluation: the lists are unchanged
    // until the next full evaluation
    fmmh_ctx* c = const_cast<fmmh_ctx*>(ctx);
    int rc = guarded(c, [&] {
      c->ensure(c->p2p_work, (1 + p2p_count_scratch(c->n_p2p)) * 8);
      launch_p2p_count((const uint2*)c->p2p.p, c->n_p2p, (const uint32_t*)c->c_count.p,
                       (unsigned long long*)c->p2p_work.p, c->s);
      c->p2p_interactions = c->read1((const unsigned long long*)c->p2p_work.p);
      c->p2p_counted = true;
    });
    if (rc) return rc;
  }
  out8[0] = ctx->n;
  out8[1] = ctx->nc;
  out8[2] = ctx->nleaves;
  out8[3] = ctx->n_m2l;
  out8[4] = ctx->n_p2p;
  out8[5] = ctx->p2p_interactions;
  out8[6] = ctx->max_level;
  out8[7] = ctx->walk_rounds;
  return FMMH_OK;
}

int fmmh_p2p_plan_counts(const fmmh_ctx* ctx, uint64_t* out2) {
  if (!ctx || !out2) return FMMH_E_STATE;
  out2[0] = ctx->p_nitems;
  out2[1] = ctx->p_nsplit;
  return FMMH_OK;
}

int fmmh_direct_sampled(size_t n, const double* pos, const double* mass, double softening, size_t nt,
                        const uint32_t* idx, double* acc, double* pot, double* f) {
  if (n == 0 || nt == 0) return FMMH_E_EMPTY;
  for (size_t t = 0; t < nt; t++)
    if (idx[t] >= n) return FMMH_E_PARAM;
  fmmh_params prm;
  fmmh_default_params(&prm);
  fmmh_ctx ctx;
  ctx.prm = prm;
  if (hipStreamCreateWithFlags(&ctx.s, hipStreamNonBlocking) != hipSuccess) return FMMH_E_HIP;
  return guarded(&ctx, [&] {
    ctx.ensure(ctx.in_pos, n * 24);
    ctx.ensure(ctx.in_mass, n * 8);
    const uint32_t nsplit = direct_split(n, nt);
    ctx.ensure(ctx.direct_ids, nt * 4);
    ctx.ensure(ctx.direct_out, nt * 40);
    ctx.ensure(ctx.direct_part, nt * (size_t)nsplit * 40);
    HCHECK(hipMemcpyAsync(ctx.in_pos.p, pos, n * 24, hipMemcpyHostToDevice, ctx.s));
    HCHECK(hipMemcpyAsync(ctx.in_mass.p, mass, n * 8, hipMemcpyHostToDevice, ctx.s));
    HCHECK(hipMemcpyAsync(ctx.direct_ids.p, idx, nt * 4, hipMemcpyHostToDevice, ctx.s));
    double* out = (double*)ctx.direct_out.p;  // acc [nt][3], pot [nt], f [nt]
    launch_direct((const double*)ctx.in_pos.p, (const double*)ctx.in_mass.p, n,
                  softening * softening, (const uint32_t*)ctx.direct_ids.p, (uint32_t)nt, nsplit,
                  (double*)ctx.direct_part.p, out, out + 3 * nt, out + 4 * nt, ctx.s);
    std::vector<double> h(nt * 5);
    HCHECK(hipMemcpyAsync(h.data(), out, nt * 40, hipMemcpyDeviceToHost, ctx.s));
    HCHECK(hipStreamSynchronize(ctx.s));
    std::memcpy(acc, h.data(), nt * 24);
    if (pot) std::memcpy(pot, h.data() + 3 * nt, nt * 8);
    if (f) std::memcpy(f, h.data() + 4 * nt, nt * 8);
  });
}

// models.rs:35-58
void fmmh_plummer(size_t n, const double center[3], uint64_t seed, double* pos, double* mass) {
  SplitMix rng{seed};
  const double m = 1.0 / (double)n;
  const double PI = 3.14159265358979323846;
  for (size_t i = 0; i < n; i++) {
    double u = rng.unit();
    double w = std::pow(u, 2.0 / 3.0);
    double r = std::sqrt(w / (1.0 - w));
    double z = 2.0 * rng.unit() - 1.0;
    double phi = 2.0 * PI * rng.unit();
    double sn = std::sqrt(1.0 - z * z);
    double sp, cp;
    sin_cos(phi, &sp, &cp);
    pos[3 * i] = center[0] + r * sn * cp;
    pos[3 * i + 1] = center[1] + r * sn * sp;
    pos[3 * i + 2] = center[2] + r * z;
    mass[i] = m;
  }
}

// Plummer sphere truncated at r <= rmax: models.rs:35-58 with the radius
// redrawn (fresh u) until it falls inside rmax; total mass 1, equal masses.
void fmmh_plummer_trunc(size_t n, const double center[3], double rmax, uint64_t seed, double* pos,
                        double* mass) {
  SplitMix rng{seed};
  const double m = 1.0 / (double)n;
  const double PI = 3.14159265358979323846;
  for (size_t i = 0; i < n; i++) {
    double r;
    do {
      double u = rng.unit();
      double w = std::pow(u, 2.0 / 3.0);
      r = std::sqrt(w / (1.0 - w));
    } while (!(r <= rmax));
    double z = 2.0 * rng.unit() - 1.0;
    double phi = 2.0 * PI * rng.unit();
    double sn = std::sqrt(1.0 - z * z);
    double sp, cp;
    sin_cos(phi, &sp, &cp);
    pos[3 * i] = center[0] + r * sn * cp;
    pos[3 * i + 1] = center[1] + r * sn * sp;
    pos[3 * i + 2] = center[2] + r * z;
    mass[i] = m;
  }
}

// models.rs:61-82
void fmmh_hernquist(size_t n, const double center[3], uint64_t seed, double* pos, double* mass) {
  SplitMix rng{seed};
  const double m = 1.0 / (double)n;
  const double PI = 3.14159265358979323846;
  for (size_t i = 0; i < n; i++) {
    double s = std::sqrt(rng.unit());
    double r = s / (1.0 - s);
    double z = 2.0 * rng.unit() - 1.0;
    double phi = 2.0 * PI * rng.unit();
    double sp = std::sqrt(1.0 - z * z);
    double sphi, cphi;
    sin_cos(phi, &sphi, &cphi);
    pos[3 * i] = center[0] + r * sp * cphi;
    pos[3 * i + 1] = center[1] + r * sp * sphi;
    pos[3 * i + 2] = center[2] + r * z;
    mass[i] = m;
  }
}

}  // extern "C"
