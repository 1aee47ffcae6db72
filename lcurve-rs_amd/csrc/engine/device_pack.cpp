// Host side of the device-packed batch (device_pack.h): CPU setup, each
// set's layout and raw grids on the setup threads, the batch-wide tables
// in one pinned block, and the pack + flux + assembly step.
#include "device_pack.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

namespace lch {

namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Byte offsets of the arrays of one block, each 256-byte aligned.
struct BlockLayout {
  size_t total = 0;
  size_t add(uint64_t bytes) {
    const size_t at = total;
    total += align256(std::max<size_t>(static_cast<size_t>(bytes), 1));
    return at;
  }
};

int thread_count(int nthreads, size_t ns) {
  int nt = nthreads;
  if (nt <= 0) {
    const char* env = std::getenv("LCURVE_HIP_THREADS");
    nt = env ? std::atoi(env) : 0;
  }
  if (nt <= 0) nt = static_cast<int>(std::thread::hardware_concurrency());
  nt = std::max(1, nt);
  return static_cast<int>(std::min<size_t>(static_cast<size_t>(nt), std::max<size_t>(ns, 1)));
}

}  // namespace

struct SetLayout {
  SetPack S;                     // scalars, segments (set-relative), lanes, per-time tables; no records
  std::vector<uint32_t> n_real;  // elements per segment
  std::vector<uint32_t> pflags;  // PSEG_* per segment
  std::vector<CodeConsts> K;     // code constants per segment
  std::vector<float> whole, none;  // the per-lane split's element cut-offs per segment
  struct Grid {
    uint32_t e0 = 0, n = 0, pos0 = 0;  // set-relative first raw element / position
    uint32_t seg_of[MAX_ECL + 1];      // set-relative segment per eclipse-pair count
  };
  std::vector<Grid> grids;             // the set's non-empty grids in segment order
  std::vector<uint32_t> seg_grid;      // per segment: its entry of `grids`
  std::vector<float4> df, pos;         // raw elements: (dirn, flux) and (posn, 0) in FP32
  std::vector<uint32_t> ecnt;          // eclipse pairs per raw element
  std::vector<double2> ecl;            // (ingress, egress) pairs, element by element
  uint64_t rec_bytes = 0;
  double setup_s = 0, layout_s = 0;
};

namespace {

// Everything of one set's device pack that the host makes: the pack's
// own segment layout, lanes and per-time tables (batch_build.cpp), the
// segments' code rules and the grids' raw elements in their own order.
std::unique_ptr<SetLayout> lay_out_set(const lc::PreparedModel& P, const lc::Data& data) {
  auto L = std::make_unique<SetLayout>();
  SetPack& S = L->S;
  set_scalars(S, P);
  const std::vector<GridRef> grids = grid_refs(P);
  std::vector<uint32_t> seg_ref;
  L->rec_bytes = lay_out_segments(S, grids, &L->n_real, &seg_ref);
  pack_lane_tables(S, P, data);
  S.n_tiles = static_cast<uint32_t>(S.lane_phi.size() / TILE);
  for (const SegDev& sg : S.segs) {
    const SegCodeRule r = seg_code_rule(sg, S.dev, P);
    L->pflags.push_back((r.cull ? PSEG_CULL : 0u) | (r.front_ok ? PSEG_FRONT : 0u) | (r.mom_ok ? PSEG_MOM : 0u) |
                        (linear_law_segment(sg, S.dev) ? PSEG_LINW : 0u));
    L->K.push_back(r.K);
    L->whole.push_back(static_cast<float>(r.thr_front + 1e-6));
    L->none.push_back(static_cast<float>(r.thr - 1e-6));
  }
  size_t ne = 0, np = 0, nq = 0;
  for (const GridRef& g : grids) {
    ne += g.pts->size();
    if (g.posn) np += g.pts->size();
    for (const lc::Point& pt : *g.pts) nq += pt.eclipse.size();
  }
  if (ne >= UINT32_MAX) throw lc::LcError("too many elements in one set");
  L->df.reserve(ne);
  L->pos.reserve(np);
  L->ecnt.reserve(ne);
  L->ecl.reserve(nq);
  std::vector<int> slot(grids.size(), -1);
  for (size_t gi = 0; gi < grids.size(); ++gi) {
    const std::vector<lc::Point>& pts = *grids[gi].pts;
    if (pts.empty()) continue;
    SetLayout::Grid g;
    g.e0 = static_cast<uint32_t>(L->df.size());
    g.n = static_cast<uint32_t>(pts.size());
    g.pos0 = grids[gi].posn ? static_cast<uint32_t>(L->pos.size()) : 0u;
    for (uint32_t& s : g.seg_of) s = UINT32_MAX;
    slot[gi] = static_cast<int>(L->grids.size());
    L->grids.push_back(g);
    for (const lc::Point& pt : pts) {
      L->df.push_back(make_float4(static_cast<float>(pt.dirn.x), static_cast<float>(pt.dirn.y),
                                  static_cast<float>(pt.dirn.z), pt.flux));
      if (grids[gi].posn)
        L->pos.push_back(make_float4(static_cast<float>(pt.posn.x), static_cast<float>(pt.posn.y),
                                     static_cast<float>(pt.posn.z), 0.0f));
      L->ecnt.push_back(static_cast<uint32_t>(pt.eclipse.size()));
      for (const auto& e : pt.eclipse) L->ecl.push_back(make_double2(e.first, e.second));
    }
  }
  for (size_t g = 0; g < S.segs.size(); ++g) {
    const int k = slot[seg_ref[g]];
    if (k < 0) throw lc::LcError("segment of an empty grid");
    L->grids[static_cast<size_t>(k)].seg_of[S.segs[g].necl] = static_cast<uint32_t>(g);
    L->seg_grid.push_back(static_cast<uint32_t>(k));
  }
  return L;
}

}  // namespace

DevicePackBatch::DevicePackBatch(Context& ctx, const std::vector<lc::Model>& models,
                                 const std::vector<lc::Data>& data, uint32_t chunk, int limb, int nthreads)
    : ctx_(ctx), models_(models), data_(data) {
  if (limb < 0 || limb > static_cast<int>(TILE)) throw std::invalid_argument("limb must be in 0..64");
  limb_ = static_cast<uint32_t>(limb);
  if (models.empty()) throw std::invalid_argument("no parameter sets");
  if (data.empty() || (data.size() != 1 && data.size() != models.size()))
    throw std::invalid_argument("give one data set, or one per parameter set");
  if (chunk < PATCH || chunk % PATCH || chunk > MAX_ITEM_PATCHES * PATCH)
    throw std::invalid_argument("chunk must be a multiple of 64 in 64..4096");
  chunk_ = chunk;
  st_.limb = limb_;
  n_times_ = data[0].size();
  for (const lc::Data& d : data)
    if (d.size() != n_times_) throw std::invalid_argument("every data set needs the same number of times");
  const size_t ns = models.size();
  std::vector<std::unique_ptr<SetLayout>> lay(ns);
  const int nt = thread_count(nthreads, ns);
  st_.threads = nt;
  const double t0 = now_s();
  std::atomic<size_t> next{0};
  std::exception_ptr fatal;
  std::mutex mu;
  auto worker = [&]() {
    for (;;) {
      const size_t s = next.fetch_add(1);
      if (s >= ns) break;
      try {
        const double a = now_s();
        std::unique_ptr<lc::PreparedModel> P;
        std::string err;
        try {
          P = lc::prepare_model(models_[s], data_of(s));
        } catch (const lc::LcError& e) {
          err = e.what();
        }
        const double b = now_s();
        std::unique_ptr<SetLayout> L;
        if (P) {
          try {
            L = lay_out_set(*P, data_of(s));
          } catch (const lc::LcError& e) {
            err = e.what();
          }
        }
        P.reset();
        if (!L) {
          L = std::make_unique<SetLayout>();
          L->S = std::move(*invalid_set(err, data_of(s)));
        }
        L->setup_s = b - a;
        L->layout_s = now_s() - b;
        lay[s] = std::move(L);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!fatal) fatal = std::current_exception();
        next.store(ns);
      }
    }
  };
  if (nt == 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
  }
  if (fatal) std::rethrow_exception(fatal);
  st_.build_wall_s = now_s() - t0;
  for (const auto& L : lay) {
    st_.setup_core_s += L->setup_s;
    st_.layout_core_s += L->layout_s;
  }
  try {
    build(std::move(lay));
    upload();
  } catch (...) {
    release();
    throw;
  }
}

DevicePackBatch::~DevicePackBatch() { release(); }

void DevicePackBatch::release_slot(Slot& s) {
  if (s.side.side) (void)hipStreamSynchronize(s.side.side);
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  for (hipEvent_t& e : s.ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (s.done) (void)hipEventDestroy(s.done);
  if (s.side.fork) (void)hipEventDestroy(s.side.fork);
  if (s.side.join) (void)hipEventDestroy(s.side.join);
  if (s.side.side) (void)hipStreamDestroy(s.side.side);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  if (s.h_counts) (void)hipHostFree(s.h_counts);
  if (s.work) (void)hipFree(s.work);
  if (s.scan_temp) (void)hipFree(s.scan_temp);
  for (void*& g : s.grow) {
    if (g) (void)hipFree(g);
    g = nullptr;
  }
  s = Slot{};
}

void DevicePackBatch::release() {
  (void)hipSetDevice(ctx_.device());
  // A device-wide wait before the frees: after a stream wait alone the
  // runtime can keep freed memory allocated (Batch::release_all).
  (void)hipDeviceSynchronize();
  for (auto& s : slots_) release_slot(*s);
  slots_.clear();
  if (h_in_) (void)hipHostFree(h_in_);
  if (d_in_) (void)hipFree(d_in_);
  h_in_ = nullptr;
  d_in_ = nullptr;
}

void DevicePackBatch::build(std::vector<std::unique_ptr<SetLayout>> lay) {
  const double t0 = now_s();
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  const size_t ns = lay.size();
  // Pass 1: sizes.
  uint64_t n_raw = 0, n_pos = 0, n_pairs = 0, n_segs = 0, n_grids = 0, n_patches = 0, n_tiles = 0;
  uint64_t n_chunks = 0, n_codes = 0, n_prows = 0, n_lanes = 0, n_times = 0, rec = 0, n_pblocks = 0, n_eclp = 0;
  uint32_t max_tiles = 0;
  bool any_iscale = false;
  for (const auto& Lp : lay) {
    const SetLayout& L = *Lp;
    n_raw += L.df.size();
    n_pos += L.pos.size();
    n_pairs += L.ecl.size();
    n_segs += L.S.segs.size();
    n_grids += L.grids.size();
    uint64_t np = 0, nch = 0;
    for (const SegDev& sg : L.S.segs) {
      np += sg.n / PATCH;
      nch += (sg.n + chunk_ - 1) / chunk_;
      if (sg.necl) n_eclp += sg.n / PATCH;
    }
    const uint32_t nt = static_cast<uint32_t>(L.S.lane_phi.size() / TILE);
    n_patches += np;
    n_tiles += nt;
    max_tiles = std::max(max_tiles, nt);
    n_chunks += nch;
    n_codes += np * nt;
    n_prows += nch * nt;
    n_lanes += L.S.lane_phi.size();
    n_times += L.S.t_lane0.size();
    rec += align256(L.rec_bytes);
    n_pblocks += (np + 255) / 256;
    if (L.S.dev.flags & SET_ISCALE) any_iscale = true;
  }
  // Sort keys are (segment << 16 | direction key) in 32 bits.
  if (n_raw >= UINT32_MAX || n_pairs >= UINT32_MAX || n_patches >= UINT32_MAX || n_tiles >= UINT32_MAX ||
      n_lanes >= UINT32_MAX || n_chunks >= UINT32_MAX || n_pblocks >= UINT32_MAX)
    throw std::invalid_argument("device-packed batch too large: split the sets");
  if (max_tiles > 65535) throw std::invalid_argument("too many lane tiles in one set");
  const uint64_t n_cand = n_chunks * max_tiles;
  if (n_cand >= UINT32_MAX) throw std::invalid_argument("too many work items: split the sets");
  if (n_codes >= UINT32_MAX) throw std::invalid_argument("too many (patch, tile) pairs: split the sets");
  uint64_t n_ktiles = 0;
  for (const auto& Lp : lay)
    for (const SetLayout::Grid& g : Lp->grids) n_ktiles += (g.n + KEY_TILE - 1) / KEY_TILE;
  st_.n_sets = ns;
  st_.n_segs = n_segs;
  st_.n_grids = n_grids;
  st_.n_elements = n_raw;
  st_.n_pairs = n_pairs;
  st_.n_patches = n_patches;
  st_.n_tiles = n_tiles;
  st_.n_chunks = n_chunks;
  st_.n_codes = n_codes;
  st_.n_cand = n_cand;
  st_.n_prows = n_prows;
  st_.n_lanes = n_lanes;
  st_.n_times = n_times;
  st_.rec_bytes = rec;
  st_.chunk = chunk_;

  // Input block.
  BlockLayout in;
  const size_t o_df = in.add(n_raw * sizeof(float4)), o_pos = in.add(n_pos * sizeof(float4));
  const size_t o_eoff = in.add((n_raw + 1) * sizeof(uint32_t)), o_ecl = in.add(n_pairs * sizeof(double2));
  const size_t o_grids = in.add(n_grids * sizeof(PackGrid)), o_psegs = in.add(n_segs * sizeof(PackSeg));
  const size_t o_gorder = in.add(n_grids * sizeof(uint32_t));
  const size_t o_kt = in.add(n_ktiles * sizeof(KeyTile));
  const size_t o_segs = in.add(n_segs * sizeof(SegDev)), o_psets = in.add(ns * sizeof(PackSet));
  const size_t o_sets = in.add(ns * sizeof(SetDev)), o_pseg = in.add(n_patches * sizeof(uint32_t));
  const size_t o_pblk = in.add(n_pblocks * sizeof(PatchBlock)), o_eclp = in.add(n_eclp * sizeof(uint32_t));
  const size_t o_chunks = in.add(n_chunks * sizeof(PackChunk)), o_tset = in.add(n_tiles * sizeof(uint32_t));
  const size_t o_cinfo = in.add(n_chunks * sizeof(uint32_t)), o_lphi = in.add(n_lanes * sizeof(double));
  const size_t o_le = in.add(n_lanes * sizeof(float4)), o_lm = in.add(n_lanes * sizeof(uint32_t));
  const size_t o_tl0 = in.add(n_times * sizeof(uint32_t)), o_tnd = in.add(n_times * sizeof(int));
  const size_t o_tsl = in.add(n_times * sizeof(double)), o_ty = in.add(n_times * sizeof(double));
  const size_t o_tfe = in.add(n_times * sizeof(double)), o_tw = in.add(n_times * sizeof(double));
  const size_t o_tph = in.add(n_times * sizeof(double));
  o_ty_ = o_ty;
  o_tfe_ = o_tfe;
  o_tw_ = o_tw;
  in_bytes_ = in.total;
  st_.input_bytes = in_bytes_;
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_in_), in_bytes_, hipHostMallocDefault), "hipHostMalloc inputs");
  hip_check(hipMalloc(reinterpret_cast<void**>(&d_in_), in_bytes_), "hipMalloc inputs");
  std::memset(h_in_, 0, in_bytes_);
  auto at = [&](size_t off) { return h_in_ + off; };
  float4* df = reinterpret_cast<float4*>(at(o_df));
  float4* pos = reinterpret_cast<float4*>(at(o_pos));
  uint32_t* eoff = reinterpret_cast<uint32_t*>(at(o_eoff));
  double2* ecl = reinterpret_cast<double2*>(at(o_ecl));
  PackGrid* pgrids = reinterpret_cast<PackGrid*>(at(o_grids));
  KeyTile* ktiles = reinterpret_cast<KeyTile*>(at(o_kt));
  uint32_t kt = 0;
  PackSeg* psegs = reinterpret_cast<PackSeg*>(at(o_psegs));
  SegDev* segs = reinterpret_cast<SegDev*>(at(o_segs));
  PackSet* psets = reinterpret_cast<PackSet*>(at(o_psets));
  SetDev* sets = reinterpret_cast<SetDev*>(at(o_sets));
  uint32_t* patch_seg = reinterpret_cast<uint32_t*>(at(o_pseg));
  PatchBlock* pblocks = reinterpret_cast<PatchBlock*>(at(o_pblk));
  uint32_t* eclp = reinterpret_cast<uint32_t*>(at(o_eclp));
  PackChunk* chunks = reinterpret_cast<PackChunk*>(at(o_chunks));
  uint32_t* tile_set = reinterpret_cast<uint32_t*>(at(o_tset));
  uint32_t* cinfo = reinterpret_cast<uint32_t*>(at(o_cinfo));
  double* lane_phi = reinterpret_cast<double*>(at(o_lphi));
  float4* lane_e = reinterpret_cast<float4*>(at(o_le));
  uint32_t* lane_meta = reinterpret_cast<uint32_t*>(at(o_lm));
  uint32_t* t_lane0 = reinterpret_cast<uint32_t*>(at(o_tl0));
  int* t_ndiv = reinterpret_cast<int*>(at(o_tnd));
  double* t_slfac = reinterpret_cast<double*>(at(o_tsl));
  double* t_y = reinterpret_cast<double*>(at(o_ty));
  double* t_ferr = reinterpret_cast<double*>(at(o_tfe));
  double* t_w = reinterpret_cast<double*>(at(o_tw));
  double* t_phase = reinterpret_cast<double*>(at(o_tph));

  // Pass 2: batch-wide tables.
  host_.resize(ns);
  seg0_.resize(ns);
  nseg_.resize(ns);
  psets_.resize(ns);
  sets_.resize(ns);
  rec0_.resize(ns);
  uint64_t e = 0, pq = 0, pr = 0, code = 0, prow = 0, recb = 0;
  uint32_t seg = 0, grid = 0, patch = 0, tile = 0, chunk = 0, lane = 0, time = 0, sorted = 0, pblk = 0, ne_cl = 0;
  uint32_t max_lanes = 0;
  eoff[0] = 0;
  for (size_t s = 0; s < ns; ++s) {
    SetLayout& L = *lay[s];
    SetPack& S = L.S;
    const uint32_t nt = static_cast<uint32_t>(S.lane_phi.size() / TILE);
    SetDev sd = S.dev;
    sd.lane_off = lane;
    sd.time_off = time;
    sd.chunk_off = chunk;
    PackSet ps{};
    ps.lane_off = lane;
    ps.n_tiles = nt;
    ps.tile0 = tile;
    ps.patch0 = patch;
    ps.code0 = code;
    ps.chunk0 = chunk;
    ps.prow_base = prow;
    host_[s] = S.host;
    seg0_[s] = seg;
    nseg_[s] = static_cast<uint32_t>(S.segs.size());
    rec0_[s] = recb;
    const uint64_t e_set = e, q_set = pq;
    for (const SetLayout::Grid& g : L.grids) {
      PackGrid G{};
      G.e0 = e_set + g.e0;
      G.n = g.n;
      for (int k = 0; k <= MAX_ECL; ++k) G.seg_of[k] = g.seg_of[k] == UINT32_MAX ? UINT32_MAX : seg + g.seg_of[k];
      for (uint32_t i0 = 0; i0 < g.n; i0 += KEY_TILE) ktiles[kt++] = KeyTile{grid, i0, std::min(KEY_TILE, g.n - i0)};
      pgrids[grid++] = G;
    }
    if (!L.df.empty()) std::memcpy(df + e, L.df.data(), L.df.size() * sizeof(float4));
    if (!L.pos.empty()) std::memcpy(pos + pq, L.pos.data(), L.pos.size() * sizeof(float4));
    for (size_t i = 0; i < L.ecnt.size(); ++i) eoff[e + i + 1] = eoff[e + i] + L.ecnt[i];
    if (!L.ecl.empty()) std::memcpy(ecl + pr, L.ecl.data(), L.ecl.size() * sizeof(double2));
    e += L.df.size();
    pq += L.pos.size();
    pr += L.ecl.size();
    uint32_t np_set = 0, nch = 0;
    for (size_t g = 0; g < S.segs.size(); ++g) {
      const SegDev& sg = S.segs[g];
      const SetLayout::Grid& gr = L.grids[L.seg_grid[g]];
      PackSeg q{};
      q.rec_off = recb + sg.rec_off;
      q.e0 = e_set + gr.e0;
      q.pos0 = static_cast<uint32_t>(q_set + gr.pos0);
      q.n_real = L.n_real[g];
      q.n_pad = sg.n;
      q.necl = sg.necl;
      q.posn = (sg.flags & SEG_POSN) ? 1u : 0u;
      q.patch0 = patch + sg.patch_off;
      q.sorted0 = sorted;
      q.set = static_cast<uint32_t>(s);
      q.flags = L.pflags[g];
      q.base_flags = sg.flags;
      q.K = L.K[g];
      q.whole = L.whole[g];
      q.none = L.none[g];
      SegDev d = sg;
      d.rec_off = q.rec_off;
      d.set = static_cast<uint32_t>(s);
      d.mom_off = q.patch0;
      psegs[seg] = q;
      segs[seg] = d;
      sorted += q.n_real;
      for (uint32_t p = 0; p < sg.n / PATCH; ++p) {
        patch_seg[q.patch0 + p] = seg;
        if (sg.necl) eclp[ne_cl++] = q.patch0 + p;
      }
      const uint32_t c = (sg.n + chunk_ - 1) / chunk_;
      for (uint32_t k = 0; k < c; ++k) {
        chunks[chunk + nch + k] = PackChunk{static_cast<uint32_t>(s), seg, k * chunk_,
                                            std::min(chunk_, sg.n - k * chunk_), nch + k};
        cinfo[chunk + nch + k] = sg.use_mask | (sg.kind << 16);
      }
      nch += c;
      np_set += sg.n / PATCH;
      ++seg;
    }
    ps.n_patches = np_set;
    ps.n_chunks = nch;
    sd.n_chunks = nch;
    sd.prow_base = prow;
    for (uint32_t p0 = 0; p0 < np_set; p0 += 256)
      pblocks[pblk++] = PatchBlock{static_cast<uint32_t>(s), patch + p0, std::min<uint32_t>(256, np_set - p0)};
    for (uint32_t t = 0; t < nt; ++t) tile_set[tile + t] = static_cast<uint32_t>(s);
    const size_t nl = S.lane_phi.size(), ntm = S.t_lane0.size();
    if (S.lane_e.size() != nl || S.lane_meta.size() != nl) throw DeviceError("lane tables out of step");
    if (S.t_ndiv.size() != ntm || S.t_slfac.size() != ntm || S.t_y.size() != ntm || S.t_ferr.size() != ntm ||
        S.t_w.size() != ntm || S.t_phase.size() != ntm)
      throw DeviceError("time tables out of step");
    std::copy(S.lane_phi.begin(), S.lane_phi.end(), lane_phi + lane);
    std::copy(S.lane_e.begin(), S.lane_e.end(), lane_e + lane);
    std::copy(S.lane_meta.begin(), S.lane_meta.end(), lane_meta + lane);
    std::copy(S.t_lane0.begin(), S.t_lane0.end(), t_lane0 + time);
    std::copy(S.t_ndiv.begin(), S.t_ndiv.end(), t_ndiv + time);
    std::copy(S.t_slfac.begin(), S.t_slfac.end(), t_slfac + time);
    std::copy(S.t_y.begin(), S.t_y.end(), t_y + time);
    std::copy(S.t_ferr.begin(), S.t_ferr.end(), t_ferr + time);
    std::copy(S.t_w.begin(), S.t_w.end(), t_w + time);
    std::copy(S.t_phase.begin(), S.t_phase.end(), t_phase + time);
    // The flux kernels' bounds, as the host-packed batch checks them.
    if ((sd.flags & SET_VALID) && sd.special0 >= sd.n_lanes) throw DeviceError("special lanes out of range");
    for (size_t t = 0; t < ntm; ++t)
      if ((sd.flags & SET_VALID) &&
          static_cast<uint64_t>(S.t_lane0[t]) + static_cast<uint64_t>(std::max(0, S.t_ndiv[t])) > sd.special0)
        throw DeviceError("time lanes out of range");
    if (sd.flags & SET_VALID) max_lanes = std::max(max_lanes, sd.special0);
    psets[s] = ps;
    sets[s] = sd;
    psets_[s] = ps;
    sets_[s] = sd;
    st_.visits += S.host.visits;
    if (S.host.valid) ++st_.n_valid;
    lane += static_cast<uint32_t>(nl);
    time += static_cast<uint32_t>(ntm);
    tile += nt;
    patch += np_set;
    chunk += nch;
    code += static_cast<uint64_t>(np_set) * nt;
    prow += static_cast<uint64_t>(nch) * nt;
    recb += align256(L.rec_bytes);
    lay[s].reset();
  }
  if (kt != n_ktiles || e != n_raw || seg != n_segs || patch != n_patches || tile != n_tiles || chunk != n_chunks || code != n_codes ||
      prow != n_prows || ne_cl != n_eclp || pblk != n_pblocks || sorted != n_raw || eoff[n_raw] != n_pairs)
    throw DeviceError("device-pack layout out of step");

  // Work block layout (each slot allocates one).
  BlockLayout w;
  WorkLayout& L = wl_;
  L.keys = w.add(n_raw * 4);
  L.vals = w.add(n_raw * 4);
  L.keys_s = w.add(n_raw * 4);
  L.vals_s = w.add(n_raw * 4);
  L.recs = w.add(rec);
  L.cones = w.add(n_patches * sizeof(Cone));
  L.moms = w.add(n_patches * MOM_STRIDE * sizeof(double));
  L.wmax = w.add(n_segs * 8);
  L.ten = w.add(n_tiles * TILE * sizeof(float4));
  L.tval = w.add(n_tiles * 8);
  L.trng = w.add(n_tiles * sizeof(double2));
  L.tcp = w.add(n_tiles * 4);
  L.codes = w.add(n_codes);
  L.efl = w.add(n_codes);
  // The per-lane split and per-chunk limb tables (limb lanes only).
  const uint64_t n_split = limb_ ? n_codes : 0, n_lch = limb_ ? n_chunks : 0;
  L.hsp = w.add(n_split);
  L.sal = w.add(n_split * 8);
  L.ssm = w.add(n_split * 8);
  L.gln = w.add(limb_ ? 4 : 0);
  L.gl = w.add(limb_ ? n_cand * 4 : 0);
  L.sst = w.add(limb_ ? 32 : 0);
  L.crev = w.add(n_lch);
  L.csh = w.add(n_lch * 4);
  L.cnq = w.add(n_lch * MAX_ITEM_PATCHES * 4);
  L.clc = w.add(n_lch * sizeof(LimbCounts));
  L.clp = w.add(n_lch * sizeof(LimbCounts));
  L.lch = w.add(n_lch * sizeof(LimbChunk));
  L.cand = w.add(n_cand * sizeof(ItemDev));
  L.cinf = w.add(n_cand * 4);
  L.cpos = w.add(n_cand * sizeof(ItemCounts));
  L.items = w.add(n_cand * sizeof(ItemDev));
  L.counts = w.add(sizeof(PackCounts));
  L.part = w.add(n_prows * TILE * sizeof(double));
  L.ltot = w.add(n_lanes * sizeof(double));
  L.flux = w.add(n_times * sizeof(double));
  L.out = w.add(ns * sizeof(SetOut));
  L.iscale = any_iscale;
  L.lcomp = w.add(any_iscale ? n_lanes * 5 * sizeof(double) : 0);
  L.comp = w.add(any_iscale ? n_times * 5 * sizeof(double) : 0);
  L.lsw = w.add(any_iscale ? n_times * 6 * sizeof(double) : 0);
  L.total = w.total;
  st_.work_bytes = L.total;
  scan_temp_bytes_ = std::max<size_t>(pack_scan_temp_bytes(n_cand, n_chunks), 16);
  n_cand_ = n_cand;
  auto din = [&](size_t off) { return d_in_ + off; };

  // The sort's workgroups take the grids largest first.
  {
    uint32_t* gorder = reinterpret_cast<uint32_t*>(at(o_gorder));
    for (uint32_t g = 0; g < n_grids; ++g) gorder[g] = g;
    std::stable_sort(gorder, gorder + n_grids, [&](uint32_t a, uint32_t b) { return pgrids[a].n > pgrids[b].n; });
  }

  // The inputs and sizes every slot shares.
  PackDev& p = pd0_;
  p.raw_df = reinterpret_cast<const float4*>(din(o_df));
  p.raw_pos = reinterpret_cast<const float4*>(din(o_pos));
  p.raw_eoff = reinterpret_cast<const uint32_t*>(din(o_eoff));
  p.raw_ecl = reinterpret_cast<const double2*>(din(o_ecl));
  p.n_raw = n_raw;
  p.grids = reinterpret_cast<const PackGrid*>(din(o_grids));
  p.n_grids = static_cast<uint32_t>(n_grids);
  p.grid_order = reinterpret_cast<const uint32_t*>(din(o_gorder));
  p.ktiles = reinterpret_cast<const KeyTile*>(din(o_kt));
  p.n_ktiles = static_cast<uint32_t>(n_ktiles);
  p.psegs = reinterpret_cast<const PackSeg*>(din(o_psegs));
  p.n_segs = static_cast<uint32_t>(n_segs);
  p.sets = reinterpret_cast<const PackSet*>(din(o_psets));
  p.n_sets = static_cast<uint32_t>(ns);
  p.patch_seg = reinterpret_cast<const uint32_t*>(din(o_pseg));
  p.n_patches = static_cast<uint32_t>(n_patches);
  p.pblocks = reinterpret_cast<const PatchBlock*>(din(o_pblk));
  p.n_pblocks = static_cast<uint32_t>(n_pblocks);
  p.ecl_patches = reinterpret_cast<const uint32_t*>(din(o_eclp));
  p.n_ecl_patches = static_cast<uint32_t>(n_eclp);
  p.chunks = reinterpret_cast<const PackChunk*>(din(o_chunks));
  p.n_chunks = static_cast<uint32_t>(n_chunks);
  p.max_tiles = max_tiles;
  p.n_tiles = static_cast<uint32_t>(n_tiles);
  p.tile_set = reinterpret_cast<const uint32_t*>(din(o_tset));
  p.lane_phi = reinterpret_cast<const double*>(din(o_lphi));
  p.lane_e = reinterpret_cast<const float4*>(din(o_le));
  p.lane_meta = reinterpret_cast<const uint32_t*>(din(o_lm));
  p.setdev = reinterpret_cast<const SetDev*>(din(o_sets));
  p.limb = limb_;

  DeviceBatch& d = d0_;
  d.n_sets = static_cast<uint32_t>(ns);
  d.max_set_lanes = max_lanes;
  d.sets = reinterpret_cast<const SetDev*>(din(o_sets));
  d.lane_phi = p.lane_phi;
  d.lane_e = p.lane_e;
  d.lane_meta = p.lane_meta;
  d.chunk_info = reinterpret_cast<const uint32_t*>(din(o_cinfo));
  d.t_lane0 = reinterpret_cast<const uint32_t*>(din(o_tl0));
  d.t_ndiv = reinterpret_cast<const int*>(din(o_tnd));
  d.t_slfac = reinterpret_cast<const double*>(din(o_tsl));
  d.t_y = reinterpret_cast<const double*>(din(o_ty));
  d.t_ferr = reinterpret_cast<const double*>(din(o_tfe));
  d.t_w = reinterpret_cast<const double*>(din(o_tw));
  d.t_phase = reinterpret_cast<const double*>(din(o_tph));
  d.n_limb_chunks = 0;
  // The segment table is written by the pack (its flags): each slot keeps
  // its own copy of the uploaded one.
  segs_off_ = o_segs;
  segs_bytes_ = n_segs * sizeof(SegDev);
  add_slot();
  st_.concat_s = now_s() - t0;
}

void DevicePackBatch::add_slot() {
  auto sp = std::make_unique<Slot>();
  Slot& s = *sp;
  try {
    hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    hip_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreateWithFlags");
    for (hipEvent_t& e : s.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipStreamCreateWithFlags(&s.side.side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    hip_check(hipEventCreateWithFlags(&s.side.fork, hipEventDisableTiming), "hipEventCreateWithFlags");
    hip_check(hipEventCreateWithFlags(&s.side.join, hipEventDisableTiming), "hipEventCreateWithFlags");
    hip_check(hipMalloc(reinterpret_cast<void**>(&s.work), wl_.total + segs_bytes_), "hipMalloc pack tables");
    hip_check(hipMalloc(&s.scan_temp, scan_temp_bytes_), "hipMalloc scan storage");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.h_counts), sizeof(PackCounts), hipHostMallocDefault),
              "hipHostMalloc counts");
    // The slot's copy of the segment table, after its work tables, from
    // the staged inputs.
    SegDev* segs = reinterpret_cast<SegDev*>(s.work + wl_.total);
    if (segs_bytes_)
      hip_check(hipMemcpyAsync(segs, h_in_ + segs_off_, segs_bytes_, hipMemcpyHostToDevice, s.stream),
                "hipMemcpyAsync segments");
    hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    const WorkLayout& L = wl_;
    auto dw = [&](size_t off) { return s.work + off; };
    s.pd = pd0_;
    PackDev& p = s.pd;
    p.segs = segs;
    p.keys = reinterpret_cast<uint32_t*>(dw(L.keys));
    p.vals = reinterpret_cast<uint32_t*>(dw(L.vals));
    p.keys_s = reinterpret_cast<uint32_t*>(dw(L.keys_s));
    p.vals_s = reinterpret_cast<uint32_t*>(dw(L.vals_s));
    p.recs = dw(L.recs);
    p.cones = reinterpret_cast<Cone*>(dw(L.cones));
    p.moms = reinterpret_cast<double*>(dw(L.moms));
    p.wmax = reinterpret_cast<uint32_t*>(dw(L.wmax));
    p.tile_en = reinterpret_cast<float4*>(dw(L.ten));
    p.tile_valid = reinterpret_cast<uint64_t*>(dw(L.tval));
    p.tile_range = reinterpret_cast<double2*>(dw(L.trng));
    p.tile_cp = reinterpret_cast<uint32_t*>(dw(L.tcp));
    p.codes = reinterpret_cast<uint8_t*>(dw(L.codes));
    p.eflags = reinterpret_cast<uint8_t*>(dw(L.efl));
    if (limb_) {
      p.has_split = reinterpret_cast<uint8_t*>(dw(L.hsp));
      p.split_all = reinterpret_cast<uint64_t*>(dw(L.sal));
      p.split_some = reinterpret_cast<uint64_t*>(dw(L.ssm));
      p.glist_n = reinterpret_cast<uint32_t*>(dw(L.gln));
      p.glist = reinterpret_cast<uint32_t*>(dw(L.gl));
      p.split_stats = reinterpret_cast<unsigned long long*>(dw(L.sst));
      p.chunk_rev = reinterpret_cast<uint8_t*>(dw(L.crev));
      p.chunk_shift = reinterpret_cast<int32_t*>(dw(L.csh));
      p.chunk_nq = reinterpret_cast<uint32_t*>(dw(L.cnq));
      p.chunk_lcnt = reinterpret_cast<LimbCounts*>(dw(L.clc));
      p.chunk_lpos = reinterpret_cast<LimbCounts*>(dw(L.clp));
      p.lchunks = reinterpret_cast<LimbChunk*>(dw(L.lch));
    }
    p.cand = reinterpret_cast<ItemDev*>(dw(L.cand));
    p.cand_info = reinterpret_cast<uint32_t*>(dw(L.cinf));
    p.cand_pos = reinterpret_cast<ItemCounts*>(dw(L.cpos));
    p.items = reinterpret_cast<ItemDev*>(dw(L.items));
    p.counts = reinterpret_cast<PackCounts*>(dw(L.counts));
    p.partials = reinterpret_cast<double*>(dw(L.part));
    s.d = d0_;
    DeviceBatch& d = s.d;
    d.segs = p.segs;
    d.items = p.items;
    d.recs = p.recs;
    d.moms = p.moms;
    if (L.iscale) {
      d.lane_comp = reinterpret_cast<double*>(dw(L.lcomp));
      d.comp = reinterpret_cast<double*>(dw(L.comp));
      d.lsw = reinterpret_cast<double*>(dw(L.lsw));
    }
    d.partials = p.partials;
    d.lane_tot = reinterpret_cast<double*>(dw(L.ltot));
    d.flux = reinterpret_cast<double*>(dw(L.flux));
    d.out = reinterpret_cast<SetOut*>(dw(L.out));
    d.limb_chunks = p.lchunks;
    // The second phase's tables, sized for one moment entry per candidate
    // to begin with (the most a pack without limb lanes writes).
    PackCounts c0{};
    c0.items.moms = static_cast<uint32_t>(n_cand_);
    ensure_capacity(s, c0);
  } catch (...) {
    release_slot(s);
    throw;
  }
  slots_.push_back(std::move(sp));
}

void DevicePackBatch::set_slots(int n) {
  if (n < 1 || n > 2) throw std::invalid_argument("slots must be 1 or 2");
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  sync_all();
  while (slots_.size() > static_cast<size_t>(n)) {
    release_slot(*slots_.back());
    slots_.pop_back();
  }
  while (slots_.size() < static_cast<size_t>(n)) add_slot();
  cur_ = last_ = 0;
}

void DevicePackBatch::sync_all() {
  for (auto& s : slots_) {
    if (s->side.side) hip_check(hipStreamSynchronize(s->side.side), "hipStreamSynchronize");
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  }
  hip_check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
}

DevicePackBatch::Slot& DevicePackBatch::next_slot() {
  Slot& s = *slots_[cur_];
  last_ = cur_;
  cur_ = (cur_ + 1) % slots_.size();
  return s;
}

void DevicePackBatch::upload() {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hipStream_t st = ctx_.stream();
  sync_all();
  const double t0 = now_s();
  hip_check(hipMemcpyAsync(d_in_, h_in_, in_bytes_, hipMemcpyHostToDevice, st), "hipMemcpyAsync inputs");
  // Each slot's copy of the segment table too.
  for (auto& sp : slots_)
    if (segs_bytes_)
      hip_check(hipMemcpyAsync(sp->pd.segs, h_in_ + segs_off_, segs_bytes_, hipMemcpyHostToDevice, st),
                "hipMemcpyAsync segments");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  st_.upload_s = now_s() - t0;
}

void DevicePackBatch::ensure_capacity(Slot& sl, const PackCounts& c) {
  PackDev& p = sl.pd;
  void** grow_ = sl.grow;
  // A table too small for `need` entries of `elem` bytes is replaced by one
  // with a quarter of headroom, so that packs of similar sets reuse it.
  // Returns the capacity in entries.
  auto grow = [&](int slot, uint64_t need, uint64_t have, size_t elem, const char* what) -> uint64_t {
    if (need <= have && grow_[slot]) return have;
    const uint64_t cap = std::max<uint64_t>(need + need / 4, 1024);
    if (grow_[slot]) {
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");  // see DevicePackBatch::release
      hip_check(hipFree(grow_[slot]), "hipFree pack table");
    }
    grow_[slot] = nullptr;
    hip_check(hipMalloc(&grow_[slot], static_cast<size_t>(cap) * elem), what);
    return cap;
  };
  const uint64_t ce = grow(0, c.items.moms, p.cap_imoms, MOM_STRIDE * sizeof(double), "hipMalloc item moments");
  (void)grow(5, c.items.moms, p.cap_imoms, sizeof(uint32_t), "hipMalloc moment entries");
  (void)grow(6, c.items.moms, p.cap_imoms, sizeof(uint64_t), "hipMalloc moment entries");
  p.cap_imoms = ce;
  p.imoms = static_cast<double*>(grow_[0]);
  p.ment_base = static_cast<uint32_t*>(grow_[5]);
  p.ment_mask = static_cast<uint64_t*>(grow_[6]);
  p.cap_masks = grow(1, c.items.masks, p.cap_masks, sizeof(uint64_t), "hipMalloc lane masks");
  p.lane_masks = static_cast<uint64_t*>(grow_[1]);
  if (limb_) {
    const uint64_t ci = grow(2, c.limb.items, p.cap_litems, sizeof(LimbItem), "hipMalloc limb items");
    (void)grow(3, c.limb.items, p.cap_litems, TILE * sizeof(uint16_t), "hipMalloc limb lane indices");
    p.cap_litems = ci;
    p.litems = static_cast<LimbItem*>(grow_[2]);
    p.lidx = static_cast<uint16_t*>(grow_[3]);
    p.cap_llanes = grow(4, c.limb.lanes, p.cap_llanes, sizeof(uint32_t), "hipMalloc limb lanes");
    p.llanes = static_cast<uint32_t*>(grow_[4]);
  }
  sl.d.imoms = p.imoms;
  sl.d.lane_masks = p.lane_masks;
  sl.d.limb_items = p.litems;
  sl.d.limb_idx = p.lidx;
  sl.d.limb_lanes = p.llanes;
}

void DevicePackBatch::pack(Slot& sl) {
  hipStream_t st = sl.stream;
  hip_check(launch_pack_count(sl.pd, sl.scan_temp, scan_temp_bytes_, st),
            "pack");
  hip_check(hipMemcpyAsync(sl.h_counts, sl.pd.counts, sizeof(PackCounts), hipMemcpyDeviceToHost, st),
            "hipMemcpyAsync counts");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  const PackCounts c = *sl.h_counts;
  const uint64_t n = static_cast<uint64_t>(c.items.a) + c.items.b + c.items.c;
  if (n > st_.n_cand || c.limb.chunks > st_.n_chunks || (!limb_ && (c.limb.chunks || c.items.masks)))
    throw DeviceError("pack counts out of range");
  // Between the phases this slot's stream is idle: its tables may be replaced.
  ensure_capacity(sl, c);
  hip_check(launch_pack_fill(sl.pd, c.items.moms, st), "pack tables");
  sl.last_counts = c;
  sl.d.n_items = static_cast<uint32_t>(n);
  sl.d.n_items_plain = c.items.a;
  sl.d.n_items_ecl = c.items.b;
  sl.d.n_limb_chunks = c.limb.chunks;
  st_.n_items = n;
  st_.n_items_plain = c.items.a;
  st_.n_items_ecl = c.items.b;
  st_.n_limb_chunks = c.limb.chunks;
  st_.n_limb_items = c.limb.items;
}

void DevicePackBatch::launch(int scale) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  Slot& sl = next_slot();
  pack(sl);
  hip_check(launch_flux_items(sl.d, sl.stream, &sl.side), "flux kernel launch");
  hip_check(launch_assemble(sl.d, scale, sl.stream), "assemble kernel launch");
  // The context stream waits for this launch: syncing it waits for every
  // launch so far, and it runs nothing itself.
  hip_check(hipEventRecord(sl.done, sl.stream), "hipEventRecord");
  hip_check(hipStreamWaitEvent(ctx_.stream(), sl.done, 0), "hipStreamWaitEvent");
}

void DevicePackBatch::run(int scale, double* chisq, double* wnok, double* sfac, double* wdwarf, double* flux) {
  launch(scale);
  sync_all();
  const Slot& sl = *slots_[last_];
  const DeviceBatch& db = sl.d;
  const PackDev& pd = sl.pd;
  last_out_.resize(sets_.size());
  hip_check(hipMemcpy(last_out_.data(), db.out, last_out_.size() * sizeof(SetOut), hipMemcpyDeviceToHost),
            "hipMemcpy out");
  for (size_t s = 0; s < last_out_.size(); ++s) {
    if (chisq) chisq[s] = last_out_[s].chisq;
    if (wnok) wnok[s] = last_out_[s].wnok;
    if (sfac) sfac[s] = last_out_[s].scale;
    if (wdwarf) wdwarf[s] = last_out_[s].wdwarf;
  }
  if (flux && st_.n_times)
    hip_check(hipMemcpy(flux, db.flux, st_.n_times * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy flux");
  // The flux pass's work, counted from the items and limb items as the
  // host-packed batch counts it.
  std::vector<ItemDev> items(db.n_items);
  if (!items.empty())
    hip_check(hipMemcpy(items.data(), pd.items, items.size() * sizeof(ItemDev), hipMemcpyDeviceToHost),
              "hipMemcpy items");
  uint64_t pe = 0, pm = 0, me = 0, pl = 0;
  for (const ItemDev& it : items) {
    pe += static_cast<uint64_t>(__builtin_popcountll(it.todo)) * PATCH * TILE;
    pm += static_cast<uint64_t>(__builtin_popcountll(it.mom) + __builtin_popcountll(it.pmom)) * PATCH * TILE;
    me += ((it.mom ? 1u : 0u) + static_cast<uint64_t>(it.ngrp)) * TILE;
  }
  if (limb_) {
    unsigned long long ss[2];
    hip_check(hipMemcpy(ss, pd.split_stats, sizeof ss, hipMemcpyDeviceToHost), "hipMemcpy split stats");
    st_.split_pairs = ss[0];
    st_.split_limb_lanes = ss[1];
  }
  if (limb_ && st_.n_chunks) {
    std::vector<LimbCounts> lc(st_.n_chunks);
    std::vector<uint32_t> nq(st_.n_chunks * MAX_ITEM_PATCHES);
    hip_check(hipMemcpy(lc.data(), pd.chunk_lcnt, lc.size() * sizeof(LimbCounts), hipMemcpyDeviceToHost),
              "hipMemcpy limb counts");
    hip_check(hipMemcpy(nq.data(), pd.chunk_nq, nq.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
              "hipMemcpy limb lanes");
    for (size_t c = 0; c < lc.size(); ++c)
      if (lc[c].chunks)
        for (uint32_t q = 0; q < MAX_ITEM_PATCHES; ++q) pl += static_cast<uint64_t>(nq[c * MAX_ITEM_PATCHES + q]) * PATCH;
  }
  st_.pairs_evaluated = pe + pl;
  st_.pairs_by_moments = pm;
  st_.moment_evals = me;
  st_.pairs_limb = pl;
}

void DevicePackBatch::set_obs(const double* flux2d, const double* ferr2d, const double* weight2d) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hipStream_t st = ctx_.stream();
  sync_all();
  const size_t n = st_.n_times * sizeof(double);
  const std::pair<const double*, size_t> obs[3] = {{flux2d, o_ty_}, {ferr2d, o_tfe_}, {weight2d, o_tw_}};
  for (const auto& o : obs) {
    if (!o.first || !n) continue;
    std::memcpy(h_in_ + o.second, o.first, n);
    hip_check(hipMemcpyAsync(d_in_ + o.second, h_in_ + o.second, n, hipMemcpyHostToDevice, st),
              "hipMemcpyAsync observations");
  }
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

void DevicePackBatch::time(int scale, int reps, double* ms_pack, double* ms_flux, double* ms_total) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  sync_all();
  double tp = 0.0, tf = 0.0;
  // The whole: `reps` launches as launch() makes them (alternating slots),
  // from the first event to the context stream's join of the last.
  hipEvent_t t0 = slots_[0]->ev[2], t1 = slots_[0]->ev[3];
  hip_check(hipEventRecord(t0, ctx_.stream()), "hipEventRecord");
  for (auto& sp : slots_) hip_check(hipStreamWaitEvent(sp->stream, t0, 0), "hipStreamWaitEvent");
  for (int r = 0; r < reps; ++r) launch(scale);
  hip_check(hipEventRecord(t1, ctx_.stream()), "hipEventRecord");
  hip_check(hipEventSynchronize(t1), "hipEventSynchronize");
  float total = 0.0f;
  hip_check(hipEventElapsedTime(&total, t0, t1), "hipEventElapsedTime");
  // The parts, one launch at a time on slot 0: the pack from its first
  // kernel to its last (the host's wait for the counts between its phases
  // included) and the flux pass.
  Slot& sl = *slots_[0];
  for (int r = 0; r < reps; ++r) {
    hip_check(hipEventRecord(sl.ev[0], sl.stream), "hipEventRecord");
    pack(sl);
    hip_check(hipEventRecord(sl.ev[1], sl.stream), "hipEventRecord");
    hip_check(launch_flux_items(sl.d, sl.stream, &sl.side), "flux kernel launch");
    hip_check(hipEventRecord(sl.ev[3], sl.stream), "hipEventRecord");
    hip_check(launch_assemble(sl.d, scale, sl.stream), "assemble kernel launch");
    hip_check(hipEventSynchronize(sl.ev[3]), "hipEventSynchronize");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, sl.ev[0], sl.ev[1]), "hipEventElapsedTime");
    tp += ms;
    hip_check(hipEventElapsedTime(&ms, sl.ev[1], sl.ev[3]), "hipEventElapsedTime");
    tf += ms;
  }
  sync_all();
  last_ = 0;
  cur_ = slots_.size() > 1 ? 1 : 0;
  if (ms_pack) *ms_pack = reps > 0 ? tp / reps : 0.0;
  if (ms_flux) *ms_flux = reps > 0 ? tf / reps : 0.0;
  if (ms_total) *ms_total = reps > 0 ? total / reps : 0.0;
}

void DevicePackBatch::check(int set, uint64_t diff[10]) {
  for (int i = 0; i < 10; ++i) diff[i] = 0;
  if (set < 0 || static_cast<size_t>(set) >= sets_.size()) throw std::invalid_argument("set index out of range");
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  sync_all();
  // The device tables of a fresh pack (on the first slot).
  Slot& sl = *slots_[0];
  pack(sl);
  hip_check(hipStreamSynchronize(sl.stream), "hipStreamSynchronize");
  const PackCounts cnt = sl.last_counts;
  const PackDev& pd = sl.pd;
  const uint64_t n_items = static_cast<uint64_t>(cnt.items.a) + cnt.items.b + cnt.items.c;
  const size_t s = static_cast<size_t>(set);
  const lc::Data& data = data_of(s);
  // The host pack of the same set.
  std::unique_ptr<lc::PreparedModel> P;
  try {
    P = lc::prepare_model(models_[s], data);
  } catch (const lc::LcError&) {
    if (sets_[s].flags & SET_VALID) diff[3] = 1;  // the device set is valid, the host's is not
    return;
  }
  std::unique_ptr<SetPack> pack;
  try {
    pack = pack_set(*P, data, false, /*elem_check=*/false);
  } catch (const lc::LcError&) {
    if (sets_[s].flags & SET_VALID) diff[3] = 1;
    return;
  }
  P.reset();
  const PackSet& ps = psets_[s];
  const uint32_t seg0 = seg0_[s], nseg = nseg_[s];
  if (pack->segs.size() != nseg || pack->n_tiles != ps.n_tiles || pack->cones.size() != ps.n_patches)
    throw DeviceError("host and device layouts of the set differ");
  auto get = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy check");
  };
  // Cones, moments and codes come from the plain records.
  std::vector<Cone> cones(ps.n_patches);
  get(cones.data(), pd.cones + ps.patch0, cones.size() * sizeof(Cone));
  for (size_t i = 0; i < cones.size(); ++i)
    if (std::memcmp(&cones[i], &pack->cones[i], sizeof(Cone)) != 0) ++diff[1];
  std::vector<double> moms(static_cast<size_t>(ps.n_patches) * MOM_STRIDE);
  get(moms.data(), pd.moms + static_cast<size_t>(ps.patch0) * MOM_STRIDE, moms.size() * sizeof(double));
  for (size_t i = 0; i < moms.size(); ++i)
    if (std::memcmp(&moms[i], &pack->moms[i], sizeof(double)) != 0) ++diff[2];
  // The device's codes are tile-major, the host's patch-major.
  std::vector<uint8_t> codes(static_cast<size_t>(ps.n_patches) * ps.n_tiles);
  get(codes.data(), pd.codes + ps.code0, codes.size());
  for (uint32_t t = 0; t < ps.n_tiles; ++t)
    for (uint32_t q = 0; q < ps.n_patches; ++q)
      if (codes[static_cast<size_t>(t) * ps.n_patches + q] != pack->ptcode[static_cast<size_t>(q) * ps.n_tiles + t])
        ++diff[3];
  // Records and segment flags after the linear-law rescaling.
  scale_linear_records(*pack);
  std::vector<char> recs(pack->recs.size());
  get(recs.data(), pd.recs + rec0_[s], recs.size());
  for (size_t i = 0; i < recs.size(); ++i)
    if (recs[i] != static_cast<char>(pack->recs[i])) ++diff[0];
  std::vector<SegDev> segs(nseg);
  get(segs.data(), pd.segs + seg0, segs.size() * sizeof(SegDev));
  for (uint32_t g = 0; g < nseg; ++g)
    if (segs[g].flags != pack->segs[g].flags) ++diff[4];
  // Work items and limb tables of a host-packed batch of this set alone,
  // at this batch's chunk and limb threshold.
  struct TuningGuard {
    Context& c;
    Tuning saved;
    ~TuningGuard() { c.tuning = saved; }
  } guard{ctx_, ctx_.tuning};
  ctx_.tuning.chunk = chunk_;
  ctx_.tuning.limb = static_cast<int>(limb_);
  std::vector<std::unique_ptr<SetPack>> one;
  one.push_back(std::move(pack));
  Batch B(ctx_, std::move(one), data.size());
  const Batch::Tables T = B.tables();
  std::vector<ItemDev> all(n_items);
  get(all.data(), pd.items, all.size() * sizeof(ItemDev));
  std::vector<ItemDev> mine;
  for (const ItemDev& it : all)
    if (it.seg >= seg0 && it.seg < seg0 + nseg) mine.push_back(it);
  diff[7] = T.items.size();
  diff[9] = T.limb_items.size();
  if (mine.size() != T.items.size()) {
    diff[5] = std::max(mine.size(), T.items.size());
  } else {
    std::vector<double> im(MOM_STRIDE);
    std::vector<uint64_t> gm;
    for (size_t i = 0; i < mine.size(); ++i) {
      ItemDev a = mine[i];
      const ItemDev& b = T.items[i];
      a.seg -= seg0;
      a.lane0 -= ps.lane_off;
      a.prow -= ps.prow_base;
      const bool same = a.seg == b.seg && a.e0 == b.e0 && a.ne == b.ne && a.lane0 == b.lane0 && a.prow == b.prow &&
                        a.todo == b.todo && a.free == b.free && a.mom == b.mom && a.pmom == b.pmom &&
                        a.ngrp == b.ngrp && a.pad == b.pad;
      if (!same) {
        ++diff[5];
        continue;
      }
      // Moment entries: the PT_MOM sum, then one per lane-mask group.
      const uint32_t ne = (a.mom ? 1u : 0u) + a.ngrp;
      for (uint32_t k = 0; k < ne; ++k) {
        get(im.data(), pd.imoms + (static_cast<size_t>(a.imom) + k) * MOM_STRIDE, MOM_STRIDE * sizeof(double));
        for (uint32_t c = 0; c < MOM_STRIDE; ++c)
          if (std::memcmp(&im[c], &T.imoms[(static_cast<size_t>(b.imom) + k) * MOM_STRIDE + c], sizeof(double)) != 0)
            ++diff[6];
      }
      if (a.ngrp) {
        gm.resize(a.ngrp);
        get(gm.data(), pd.lane_masks + a.lm_off, gm.size() * sizeof(uint64_t));
        for (uint32_t k = 0; k < a.ngrp; ++k)
          if (gm[k] != T.lane_masks[static_cast<size_t>(b.lm_off) + k]) ++diff[6];
      }
    }
  }
  // Limb chunks of this set, in order, with their items, lane indices and
  // lane lists.
  std::vector<LimbChunk> lch(cnt.limb.chunks), lmine;
  get(lch.data(), pd.lchunks, lch.size() * sizeof(LimbChunk));
  for (const LimbChunk& c : lch)
    if (c.seg >= seg0 && c.seg < seg0 + nseg) lmine.push_back(c);
  if (lmine.size() != T.limb_chunks.size()) {
    diff[8] = std::max(lmine.size(), T.limb_chunks.size());
    return;
  }
  for (size_t i = 0; i < lmine.size(); ++i) {
    const LimbChunk& a = lmine[i];
    const LimbChunk& b = T.limb_chunks[i];
    if (a.seg - seg0 != b.seg || a.chunk != b.chunk || a.n_items != b.n_items || a.n_lanes != b.n_lanes ||
        a.shift != b.shift) {
      ++diff[8];
      continue;
    }
    std::vector<LimbItem> li(a.n_items);
    std::vector<uint16_t> lx(static_cast<size_t>(a.n_items) * TILE);
    std::vector<uint32_t> ll(a.n_lanes);
    get(li.data(), pd.litems + a.item0, li.size() * sizeof(LimbItem));
    get(lx.data(), pd.lidx + static_cast<size_t>(a.item0) * TILE, lx.size() * sizeof(uint16_t));
    get(ll.data(), pd.llanes + a.lane_off, ll.size() * sizeof(uint32_t));
    for (uint32_t k = 0; k < a.n_items; ++k) {
      const LimbItem& h = T.limb_items[b.item0 + k];
      if (li[k].e0 != h.e0 || li[k].split != h.split) ++diff[8];
      for (uint32_t j = 0; j < TILE; ++j)
        if (lx[static_cast<size_t>(k) * TILE + j] != T.limb_idx[(static_cast<size_t>(b.item0) + k) * TILE + j])
          ++diff[8];
    }
    for (uint32_t u = 0; u < a.n_lanes; ++u)
      if (ll[u] - ps.lane_off != T.limb_lanes[b.lane_off + u]) ++diff[8];
  }
}

}  // namespace lch
