// Work lists and launches of the wave stage kernels (wave_stage_kernels.h):
// the interior box, absorbing-band and free-surface tiles of one RK stage,
// the XCD-contiguous block numbering, and the per-stage launch.
#include "wave_stage_kernels.h"

namespace cgfdm {
using namespace wave;

namespace {

template <class MT>
WaveArgs<MT> make_args(Ctx& c, const StageBufs& b, int irk, bool rhs_only, float* rhs_out) {
  WaveArgs<MT> a{};
  a.Win = b.Win; a.Wout = b.Wout; a.S = b.S; a.T = b.T;
  a.met = (const MT*)c.met;
  a.slabH = c.slabH;
  a.rhs = rhs_out;
  a.ttop = c.trac_top;
  a.itop = c.img_top;
  a.cold = (const WaveCold*)c.wave_cold;
  bool any = false;
  for (int f = 0; f < 6; f++) any |= c.pml.on[f] != 0;
  a.pml = any ? 1 : 0;
  a.ni = c.d.ni; a.nj = c.d.nj; a.nk = c.d.nk; a.ny = c.d.ny; a.ii0 = c.d.i0 - 3;
  a.nyz = (unsigned)c.d.pyz;
  a.py = c.d.py; a.jpad = c.d.jpad;
  a.rDH = (float)(1.0 / c.p.DH);
  for (int m = 0; m < 10; m++) a.mref[m] = c.met_ref[m];
  a.pal = std::is_same<MT, MetBf16Delta>::value ? (const float4*)c.met_pal : nullptr;
  a.npal = c.met_npal;
  a.met_skip = std::is_same<MT, MetBf16Delta>::value ? c.met_skip : 0u;
  const double RA[4] = {0.5, 0.5, 1.0, 1e10}, RB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  a.rka = (float)(RA[irk] * c.p.DT);
  a.rkb = (float)(RB[irk] * c.p.DT);
  a.irk = irk;
  a.freenode = c.p.FreeSurface ? 1 : 0;
  a.cerjan = c.p.useCerjan ? 1 : 0;
  a.fault = c.p.Rupture ? 1 : 0;
  a.rhs_only = rhs_only ? 1 : 0;
  a.aux_fused = c.pml_fused ? 1 : 0;
  return a;
}

// Tiles [lo, hi) of one axis: the edge band below, tiles of length t
// starting at the interior's first index, the edge band above.  Each
// segment is (start, length, interior?).
// A tile range along one axis: lanes [0, n) of a tile starting at index s,
// of which [o, n) are active.  Low-side band segments end at the interior
// edge, so a partial tile's idle lanes fall on the ghost side (their clamped
// loads all hit the ghost column) rather than on interior points.
struct Seg { int s, n; bool in; int o; };
std::vector<Seg> axis_segments(int n, int lo, int hi, int t) {
  std::vector<Seg> v;
  const int a = std::min(lo, n), b = std::max(a, n - hi);
  const int m = (a + t - 1) / t;
  for (int q = 0; q < m; q++) {
    const int x = a - (m - q) * t;
    v.push_back({x, t, false, std::max(0, -x)});
  }
  for (int x = a; x < b; x += t) v.push_back({x, std::min(t, b - x), true, 0});
  for (int x = b; x < n; x += t) v.push_back({x, std::min(t, n - x), false, 0});
  return v;
}

}  // namespace

const char* const kProfNames[PK_COUNT] = {"stage_interior", "stage_band", "stage_surface", "pml_aux",
                                          "fault_pre", "fault_deriv", "fault_post", "rup_sensor", "fault_rs"};

// First interior column along j: the width of the low absorbing band, or
// with WaveBandRound that width rounded up to whole TJ-wide tiles, so that
// the band's tiles are full width (a 32-wide tile holding 20 band columns
// pays the j halo in whole cache lines for few points, in a launch of its
// own that cannot share those lines with the interior tile beside it).  The
// columns between the band and this edge are then plain points, which the
// band kernel treats like the interior one.
int interior_lo_j(const Params& p) {
  const int band = std::max(p.usePML ? p.PML_N : 0, p.useCerjan ? p.DAMP_N : 0);
  if (p.WaveBandRound == 0) return band;
  return (band + TJ - 1) / TJ * TJ;
}

// Uploads the off-hot-path pointer table and builds the stage-kernel work
// lists.  The interior box excludes every face band where a PML profile or
// the Cerjan sponge is non-trivial and the three free-surface rows; its
// tiles start at the box edge so that no interior tile straddles a band.
// Interior items come first (chunks sized for about 4096 blocks in total, so
// 256 CUs stay balanced), then the edge items; each list is chunk-major so that
// one XCD's contiguous share covers neighbouring tiles.
int upload_wave_cold(Ctx& c) {
  WaveCold w{};
  for (int x = 0; x < 3; x++) { w.pA[x] = c.pml.A[x]; w.pB[x] = c.pml.B[x]; w.pD[x] = c.pml.D[x]; w.cj[x] = c.cj[x]; }
  for (int f = 0; f < 6; f++) { w.auxY[f] = c.pml.Y[f]; w.auxS[f] = c.pml.S[f]; w.auxT[f] = c.pml.T[f]; w.on[f] = c.pml.on[f]; }
  w.mVx = c.matVx2Vz; w.mVy = c.matVy2Vz;
  w.N = c.pml.N;
  if (!c.wave_cold) {
    if (hipMalloc(&c.wave_cold, sizeof(WaveCold)) != hipSuccess) return 1;
    c.allocations.push_back(c.wave_cold);
  }
  if (hipMemcpy(c.wave_cold, &w, sizeof w, hipMemcpyHostToDevice) != hipSuccess) return 1;
  if (c.p.FreeSurface && !c.trac_top) {  // the image stencil reaches six rows down (NZ >= 7)
    if (hipMalloc(&c.trac_top, sizeof(float) * 9 * 6 * (size_t)c.d.nx * c.d.ny) != hipSuccess) return 1;
    c.allocations.push_back(c.trac_top);
    if (hipMalloc(&c.img_top, sizeof(float) * 9 * (size_t)c.d.ni * c.d.nj) != hipSuccess) return 1;
    c.allocations.push_back(c.img_top);
  }

  const Params& p = c.p;
  const int n3[3] = {c.d.ni, c.d.nj, c.d.nk};
  const bool freenode = p.FreeSurface != 0;
  int lo[3], hi[3];
  for (int ax = 0; ax < 3; ax++) {
    lo[ax] = ax == 1 ? interior_lo_j(p) : std::max(c.pml.on[2 * ax] ? c.pml.N : 0, p.useCerjan ? p.DAMP_N : 0);
    hi[ax] = std::max(c.pml.on[2 * ax + 1] ? c.pml.N : 0, p.useCerjan && (ax < 2 || !freenode) ? p.DAMP_N : 0);
  }
  // WaveSurfaceTiles: the free-surface tiles are exactly the top TK rows
  // (full tiles) and the interior's k remainder joins the low band instead;
  // their j tiles are laid out on their own, line-aligned and full width
  const bool surf = freenode && p.WaveSurfaceTiles != 0;
  if (freenode) hi[2] = std::max(hi[2], surf ? TK : 3);
  // the interior box holds whole tiles along j and k: a remainder joins the
  // high-side band (or the free-surface tiles), whose kernels treat any point
  // (WaveTallInterior: interior tiles of 2 TK rows in blocks of 2 NTHR
  // threads, which re-read the k halo rows for twice as many points)
  const bool tall = p.WaveTallInterior != 0;
  c.wave_tall = tall;
  // WaveInteriorKernel: 0 the general kernel on the interior tiles, 1 the
  // pipelined interior kernel (64 x 8 tiles, LDS plane ring, one block per
  // CU), 2 the pipelined kernel on 64 x 6 tiles (384 threads, two blocks per
  // CU), 3 on 64 x 4 tiles (256 threads, three blocks per CU), 4 on 64 x 8
  // tiles with an LDS double buffer instead of the ring
  // WaveInteriorKernel -1 (default) picks 3 when the interior launch holds
  // at least 16 rounds of resident 64 x 4 blocks (three per CU; fewer rounds
  // leave the last ones ragged: 1600 x 1000 x 500, 88 k blocks: 90-95 ms of
  // interior per step against 96-101 for 1; 400 x 300 x 200, 2 k blocks:
  // 2.67 against 2.49 ms), else 1
  int wik = p.WaveInteriorKernel;
  if (wik < 0) {
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    const long ej = std::max(0, n3[1] - lo[1] - hi[1]), ek = std::max(0, n3[2] - lo[2] - hi[2]);
    const long span = std::max(0, n3[0] - lo[0] - hi[0]);
    const long blocks4 = (ej / TJ) * (ek / TK4) * ((span + 31) / 32);
    wik = blocks4 >= 16L * 3 * cus ? 3 : 1;
  }
  c.wave_pipe = !tall && wik >= 1;
  c.wave_ring = wik == 1;
  // WaveBandKernel 1: the pipelined kernel on the band tiles too, which adds
  // the aux terms of one face per axis: used when the two ADE-PML bands of
  // every axis are disjoint (2 N <= the axis length)
  bool disjoint = true;
  for (int ax = 0; ax < 3; ax++)
    if (c.pml.on[2 * ax] && c.pml.on[2 * ax + 1] && 2 * c.pml.N > (ax == 0 ? c.d.ni : ax == 1 ? c.d.nj : c.d.nk))
      disjoint = false;
  // -1 (automatic): 64 x 4 band tiles beside the 64 x 4 interior tiles of a
  // large grid, 64 x 8 ones on a grid too small for those (400 x 300 x 200:
  // bands 3.28 against 3.35 ms per step)
  const int wbk = p.WaveBandKernel >= 0 ? p.WaveBandKernel : wik == 3 ? 2 : 1;
  c.wave_band_pipe = wbk >= 1 && disjoint;
  c.wave_band_tk = c.wave_band_pipe && wbk == 2 ? TK4 : TK;  // 2: 64 x 4 band tiles, two blocks per CU
  // WaveSurfaceKernel 1: the pipelined kernel on the free-surface tiles too
  // (with the band terms; their auxiliary fields stay with pml_aux_kernel)
  c.wave_surf_pipe = disjoint && p.WaveSurfaceKernel >= 1;
  // PmlAuxFused 1: the pipelined band tiles advance the auxiliary fields of
  // their points on one face only; pml_aux_kernel keeps the edges of the
  // absorbing box (points on two or three faces, whose words the band tiles
  // prefetch with the first face's but do not advance) and the free-surface
  // rows (1600 x 1000 x 500: 138.5 -> 128.6 ms per step; 400 x 300 x 200:
  // 7.67 -> 6.90 ms; interleaved runs)
  c.pml_fused = c.wave_band_pipe && p.PmlAuxFused != 0;
  const int tki = tall ? 2 * TK : (c.wave_pipe && wik == 2) ? TK6 : (c.wave_pipe && wik == 3) ? TK4 : TK;  // interior tile rows
  c.wave_tki = tki;
  for (int ax = 1; ax < 3; ax++) {
    const int t = ax == 1 ? TJ : tki, e = n3[ax] - lo[ax] - hi[ax];
    if (e > t) (ax == 2 && surf ? lo[ax] : hi[ax]) += e % t;
  }
  const std::vector<Seg> sj = axis_segments(n3[1], lo[1], hi[1], TJ);
  const std::vector<Seg> sk = axis_segments(n3[2], lo[2], hi[2], TK);
  const std::vector<Seg> ski = tki != TK ? axis_segments(n3[2], lo[2], hi[2], tki) : sk;  // interior tiles' rows
  const std::vector<Seg>& sji = sj;
  const int nk = n3[2];
  auto is_top = [&](const Seg& b) { return freenode && b.s + b.n > nk - 3; };
  std::vector<Seg> sjt = sj;  // j tiles of the free-surface rows
  if (surf) {
    sjt.clear();
    const int s0 = lo[1] - 32 * ((lo[1] + 31) / 32);  // j = s0 starts a 128-byte line
    for (int x = s0; x < n3[1]; x += TJ) sjt.push_back({x, std::min(TJ, n3[1] - x), false, std::max(0, -x)});
  }
  int tiles[3] = {0, 0, 0};
  for (const Seg& b : sk)
    for (const Seg& a : sj)
      if (!is_top(b)) tiles[a.in && b.in ? 0 : 1]++;
  for (const Seg& b : sk)
    if (is_top(b)) tiles[2] += (int)sjt.size();
  int tiles_in = 0;  // interior blocks per chunk
  for (const Seg& b : ski)
    for (const Seg& a : sji) tiles_in += a.in && b.in;
  const int ni = n3[0], ia = std::min(lo[0], ni), ib = std::max(ia, ni - hi[0]);
  // planes per block of each kind: short chunks (several thousand blocks per
  // launch) balance the CUs better than the longer chunks would save on the
  // 4-plane warm-up of the x-queue; the few free-surface tiles take at least
  // 4 planes too (shorter chunks re-read the warm-up planes for 1-2 planes
  // of work; 0.58 against 0.52 ms per step on 400 x 300 x 200)
  auto chunk = [&](int span, int ntiles, int target, int minlen) {
    span = std::max(span, 1);
    const int chunks = std::max(1, std::min(span, (target + std::max(ntiles, 1) - 1) / std::max(ntiles, 1)));
    return std::max((span + chunks - 1) / chunks, std::min(minlen, span));
  };
  // The pipelined interior kernel streams 32 planes per block: its loads
  // overlap the compute, so a longer chunk mostly saves warm-up planes
  // (1600 x 1000 x 500: 97.7 ms of interior per step against 106.2 for the
  // ~300-plane chunks of a 4096-block target and 104.5 for 16 planes;
  // 400 x 300 x 200: 2.44 against 2.70 ms for ~8 planes).
  const int ich_in = p.WaveIChunk > 0 ? p.WaveIChunk
                     : c.wave_pipe   ? std::min(std::max(ib - ia, 1), 32)
                                     : chunk(ib - ia, tiles_in, tall ? 2048 : 4096, 4);
  // band tiles stream 20 planes (1600 x 1000 x 500: 23.4 ms per step against
  // 24.6, 25.0 and 32.9 for 16, 32 and 48 planes and 29.6 for the ~4 planes
  // of an 8192-block target; 400 x 300 x 200: 2.55 against 2.68 ms)
  const int ich_band = p.WaveBandIChunk > 0 ? p.WaveBandIChunk : std::min(std::max(ni, 1), 20);
  const int ich_top = p.WaveTopIChunk > 0 ? p.WaveTopIChunk : chunk(ni, tiles[2], 2048, 4);
  // lists: 0 interior, 1 absorbing bands (64-wide tiles), 2 absorbing bands
  // of at most 32 columns along j (32 x 16 tiles, half the idle lanes),
  // 3 tiles holding free-surface rows
  std::vector<int4> lists[4];
  for (int kd = 0; kd < 3; kd++) c.wave_points[kd] = 0;
  for (int kd = 0; kd < 3; kd++) c.aux_read_pairs[kd] = c.aux_update_pairs[kd] = 0;
  // points of [lo, hi) along axis ax inside that axis's absorbing bands
  auto in_bands = [&](int ax, int lo, int hi) {
    const int N = c.pml.N, nn = n3[ax];
    const int a0 = 0, a1 = c.pml.on[2 * ax] ? std::min(N, nn) : 0;
    const int b0 = c.pml.on[2 * ax + 1] ? std::max(nn - N, 0) : nn, b1 = nn;
    auto len = [&](int x0, int x1) { return std::max(0, std::min(x1, hi) - std::max(x0, lo)); };
    return len(a0, a1) + len(b0, b1) - len(std::max(a0, b0), std::min(a1, b1));
  };
  auto push = [&](int list, const Seg& a, const Seg& b, int i0, int i1) {
    lists[list].push_back(make_int4(a.s, b.s, i0, i1));
    lists[list].push_back(make_int4(a.n, b.n, a.o, b.o));
    const int kd = list == 0 ? 0 : list == 3 ? 2 : 1;
    const long Li = i1 - i0, Lj = a.n - a.o, Lk = b.n - b.o;
    c.wave_points[kd] += Li * Lj * Lk;
    const long bi = in_bands(0, i0, i1), bj = in_bands(1, a.s + a.o, a.s + a.n), bk = in_bands(2, b.s + b.o, b.s + b.n);
    c.aux_read_pairs[kd] += bi * Lj * Lk + Li * bj * Lk + Li * Lj * bk;
    // the fused update in the band tiles: points on one face only
    if (kd == 1 && c.pml_fused)
      c.aux_update_pairs[kd] += bi * (Lj - bj) * (Lk - bk) + (Li - bi) * bj * (Lk - bk) + (Li - bi) * (Lj - bj) * bk;
  };
  const auto narrow = [&](const Seg& a) { return !a.in && a.n - a.o <= TJN; };
  // bands of at most TJB columns get tiles exactly TJB wide (with the
  // pipelined band kernel), so that a 20-column band keeps every lane busy
  // (a 32-wide tile idles 12 of its 32; 1600 x 1000 x 500: 2.8 ms per stage
  // for the 32 M points of the two j bands against 2.4 ms for the 48 M
  // points of the 64-wide band tiles); wider narrow bands get TJN x TKN tiles
  int wmax = 0;
  for (const Seg& a : sj)
    if (narrow(a)) wmax = std::max(wmax, a.n - a.o);
  const bool exact = c.wave_band_pipe && wmax > 0 && wmax <= TJB;
  const int tjn = exact ? TJB : TJN, tkn = exact ? TKB : TKN;
  c.wave_tjn = tjn;
  // the same active columns in a tjn-wide tile: anchored at the interior
  // edge on the low side, at the band start on the high side
  const auto narrowed = [&](const Seg& a) {
    const int w = a.n - a.o;
    return a.o > 0 ? Seg{a.s + a.n - tjn, tjn, false, tjn - w} : Seg{a.s, w, false, 0};
  };
  int ktop = nk;  // first row of the free-surface tiles
  for (const Seg& b : sk)
    if (is_top(b)) ktop = std::min(ktop, b.s);
  c.wave_ktop = ktop;
  for (int i0 = ia; i0 < ib; i0 += ich_in)
    for (const Seg& b : ski) for (const Seg& a : sji) if (a.in && b.in) push(0, a, b, i0, std::min(ib, i0 + ich_in));
  // band tiles of TK rows, or of TK4 rows with WaveBandKernel 2 (the rows
  // from ktop up are the free-surface tiles')
  const std::vector<Seg> skb = c.wave_band_tk == TK4 ? axis_segments(n3[2], lo[2], hi[2], TK4) : sk;
  for (int i0 = 0; i0 < ni; i0 += ich_band)
    for (const Seg& b : skb)
      for (const Seg& a : sj) {
        const int i1 = std::min(ni, i0 + ich_band);
        if (b.s >= ktop || narrow(a)) continue;
        if (!(a.in && b.in)) { push(1, a, b, i0, i1); continue; }
        // interior tile: only the parts of [i0, i1) outside [ia, ib)
        if (i0 < ia) push(1, a, b, i0, std::min(i1, ia));
        if (i1 > ib) push(1, a, b, std::max(i0, ib), i1);
      }
  for (int i0 = 0; i0 < ni; i0 += ich_band)
    for (int k0 = 0; k0 < ktop; k0 += tkn)
      for (const Seg& a : sj)
        if (narrow(a)) push(2, narrowed(a), Seg{k0, std::min(tkn, ktop - k0), false, 0}, i0, std::min(ni, i0 + ich_band));
  for (int i0 = 0; i0 < ni; i0 += ich_top)
    for (const Seg& b : sk)
      if (is_top(b))
        for (const Seg& a : sjt) push(3, a, b, i0, std::min(ni, i0 + ich_top));
  // boxes of the auxiliary-update kernel: every absorbing face whole; with
  // the fused update only the points the band tiles leave to it, the
  // free-surface rows and the edges of the absorbing box (points on two or
  // three faces).  In a face's cross-section (u, v) the edges are the
  // u-band rows for every v, then the v-band columns of the u-middle; the
  // free-surface rows add the u-middle x v-middle rows from ktop up.
  {
    const int N = c.pml.N, nn3[3] = {c.d.ni, c.d.nj, c.d.nk};
    auto bands = [&](int axis) {
      std::vector<std::pair<int, int>> v;  // (start, count)
      if (c.pml.on[2 * axis]) v.push_back({0, N});
      if (c.pml.on[2 * axis + 1]) v.push_back({nn3[axis] - N, N});
      return v;
    };
    auto middle = [&](int axis) {
      const int lo = c.pml.on[2 * axis] ? N : 0, hi = c.pml.on[2 * axis + 1] ? nn3[axis] - N : nn3[axis];
      return std::make_pair(lo, std::max(hi - lo, 0));
    };
    std::vector<int> bx;
    int blocks = 0;
    c.aux_update_pairs[3] = 0;
    // box in face-local (j, k, i) coordinates
    auto add = [&](int f, int j0, int nj_, int k0, int nk_, int i0, int ni_) {
      if (nj_ <= 0 || nk_ <= 0 || ni_ <= 0) return;
      bx.insert(bx.end(), {f, j0, nj_, k0, nk_, i0, ni_, blocks});
      blocks += (int)(((long)nj_ * nk_ * ni_ + 255) / 256);
      c.aux_update_pairs[3] += (long)nj_ * nk_ * ni_;
    };
    for (int f = 0; f < 6; f++) {
      if (!c.pml.on[f]) continue;
      const int ax = f / 2;
      const int pnj = ax == 1 ? N : c.d.nj, pnk = ax == 2 ? N : c.d.nk, pni = ax == 0 ? N : c.d.ni;
      if (!c.pml_fused) { add(f, 0, pnj, 0, pnk, 0, pni); continue; }
      if (ax == 0) {  // u = j, v = k
        const auto mu = middle(1), mv = middle(2);
        for (auto b : bands(1)) add(f, b.first, b.second, 0, pnk, 0, pni);
        for (auto b : bands(2)) add(f, mu.first, mu.second, b.first, b.second, 0, pni);
        const int s0 = std::max(ktop, mv.first);
        add(f, mu.first, mu.second, s0, mv.first + mv.second - s0, 0, pni);
      } else if (ax == 1) {  // u = i, v = k
        const auto mu = middle(0), mv = middle(2);
        for (auto b : bands(0)) add(f, 0, pnj, 0, pnk, b.first, b.second);
        for (auto b : bands(2)) add(f, 0, pnj, b.first, b.second, mu.first, mu.second);
        const int s0 = std::max(ktop, mv.first);
        add(f, 0, pnj, s0, mv.first + mv.second - s0, mu.first, mu.second);
      } else {  // u = i, v = j; the face's own k: free-surface rows inside the band
        const auto mu = middle(0), mv = middle(1);
        for (auto b : bands(0)) add(f, 0, pnj, 0, pnk, b.first, b.second);
        for (auto b : bands(1)) add(f, b.first, b.second, 0, pnk, mu.first, mu.second);
        const int off = (f & 1) ? c.d.nk - N : 0, s0 = std::max(ktop - off, 0);
        add(f, mv.first, mv.second, s0, N - s0, mu.first, mu.second);
      }
    }
    c.pml_nbox = (int)bx.size() / kPmlBoxInts;
    c.pml_blocks = blocks;
    if (c.pml_boxes) { (void)hipFree(c.pml_boxes); c.pml_boxes = nullptr; }
    if (!bx.empty()) {
      if (hipMalloc(&c.pml_boxes, bx.size() * sizeof(int)) != hipSuccess) return 1;
      if (hipMemcpy(c.pml_boxes, bx.data(), bx.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return 1;
    }
  }
  std::vector<int4> all;
  for (int l = 0; l < 4; l++) {
    all.insert(all.end(), lists[l].begin(), lists[l].end());
    c.wave_nkind[l] = (int)lists[l].size() / 2;
  }
  if (c.wave_work) { (void)hipFree(c.wave_work); c.wave_work = nullptr; }
  if (hipMalloc(&c.wave_work, all.size() * sizeof(int4)) != hipSuccess) return 1;
  if (hipMemcpy(c.wave_work, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess) return 1;
  c.wave_nwork = (int)all.size() / 2;
  return 0;
}

void launch_wave_stage(Ctx& c, const StageBufs& b, int irk, int fx, int fy, int fz, bool rhs_only, float* rhs_out) {
  if (c.p.MetricBF16 == 2) {
    WaveArgs<MetBf16Delta> a = make_args<MetBf16Delta>(c, b, irk, rhs_only, rhs_out);
    launch_wave_dispatch(a, c, fx, fy, fz, c.stream);
  } else if (c.p.MetricBF16) {
    WaveArgs<uint16_t> a = make_args<uint16_t>(c, b, irk, rhs_only, rhs_out);
    launch_wave_dispatch(a, c, fx, fy, fz, c.stream);
  } else {
    WaveArgs<float> a = make_args<float>(c, b, irk, rhs_only, rhs_out);
    launch_wave_dispatch(a, c, fx, fy, fz, c.stream);
  }
}

}  // namespace cgfdm
