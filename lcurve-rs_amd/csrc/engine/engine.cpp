// Host side of the flux engine: packs CPU-setup output into the HBM layout
// of layout.h, builds work items, launches (optionally through a captured
// HIP graph), times, autotunes and reads back results.
#include "engine.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "batch_build.h"
#include "scratch.h"

namespace lch {

namespace {
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }
}  // namespace

// ----------------------------------------------------------------- context
Context::Context(int device) : device_(device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw DeviceError("no HIP device available");
  if (device < 0 || device >= n)
    throw DeviceError("device id " + std::to_string(device) + " out of range (" + std::to_string(n) +
                      " devices)");
  hip_check(hipSetDevice(device), "hipSetDevice");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  cu_count_ = prop.multiProcessorCount;
  // Default priority on purpose: streams of other priorities are avoided.
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  // The static scratch's address on this device, with the device current.
  scratch_ = scratch_device_ptr(device, &scratch_err_);
}

Context::~Context() {
  if (stream_) {
    (void)hipSetDevice(device_);
    (void)hipStreamDestroy(stream_);
  }
}

// ------------------------------------------------------------------- batch
template <class T>
T* Batch::dalloc(size_t n) {
  size_t bytes = std::max<size_t>(n * sizeof(T), 16);
  if (mem_.arena) return static_cast<T*>(item_alloc(bytes, "batch tables"));
  void* p = nullptr;
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  allocs_.push_back(p);
  return static_cast<T*>(p);
}

void Batch::dfree_all() {
  for (void* p : allocs_) (void)hipFree(p);
  allocs_.clear();
}

void* Batch::item_alloc(size_t bytes, const char* what) {
  bytes = std::max<size_t>(bytes, 16);
  if (mem_.arena) {
    const size_t at = align256(arena_used_);
    if (at + bytes > mem_.arena_bytes)
      throw ArenaFull(std::string(what) + ": the call needs more than the " +
                              std::to_string(mem_.arena_bytes) + " bytes of device memory it was given");
    arena_used_ = at + bytes;
    return mem_.arena + at;
  }
  void* p = nullptr;
  hip_check(hipMalloc(&p, bytes), (std::string("hipMalloc ") + what).c_str());
  return p;
}

void Batch::item_free(void* p) {
  if (p && !mem_.arena) (void)hipFree(p);
}

Batch::Batch(Context& ctx, std::vector<std::unique_ptr<SetPack>> packs, size_t n_times,
             const BatchMemory* mem)
    : ctx_(ctx), n_times_(n_times) {
  if (mem) mem_ = *mem;
  if (mem_.arena) use_graph = false;
  // A failure part-way (an allocation, an upload) frees what was created:
  // the destructor does not run for a constructor that throws.
  try {
    init(std::move(packs));
  } catch (...) {
    release_all();
    throw;
  }
}

void Batch::init(std::vector<std::unique_ptr<SetPack>> packs) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  const size_t ns = packs.size();
  host_.resize(ns);
  sets_.resize(ns);
  pt_base_.assign(ns, 0);
  pt_tiles_.assign(ns, 0);
  lsplit_.resize(ns);

  // Global offsets.
  uint64_t rec_total = 0;
  size_t nseg = 0, nlane = 0, ntime = 0;
  std::vector<uint64_t> rec_base(ns);
  for (size_t s = 0; s < ns; ++s) {
    rec_base[s] = rec_total;
    rec_total += (packs[s]->recs.size() + 255) / 256 * 256;
    nseg += packs[s]->segs.size();
    nlane += packs[s]->lane_phi.size();
    ntime += packs[s]->t_lane0.size();
  }
  if (mem_.recs) {
    if (ns != 1 || !packs[0]->recs.empty()) throw DeviceError("resident records need a one-set batch without records");
    rec_total = mem_.rec_bytes;
  }
  rec_bytes_ = rec_total;
  n_lanes_ = nlane;

  char* d_recs = mem_.recs ? const_cast<char*>(mem_.recs) : dalloc<char>(rec_total);
  segs_.reserve(nseg);
  lane_phi_.reserve(nlane);
  lane_meta_.reserve(nlane);
  t_lane0_.reserve(ntime);
  t_ndiv_.reserve(ntime);
  t_slfac_.reserve(ntime);
  t_y_.reserve(ntime);
  t_ferr_.reserve(ntime);
  t_w_.reserve(ntime);

  for (size_t s = 0; s < ns; ++s) {
    SetPack& P = *packs[s];
    SetDev sd = P.dev;
    sd.lane_off = static_cast<uint32_t>(lane_phi_.size());
    sd.time_off = static_cast<uint32_t>(t_lane0_.size());
    sets_[s] = sd;
    host_[s] = P.host;
    SetHost& H = host_[s];
    if ((sd.flags & SET_VALID) && P.n_tiles != P.lane_phi.size() / TILE) throw DeviceError("patch codes out of step");
    pt_base_[s] = ptcode_.size();
    pt_tiles_[s] = P.n_tiles;
    ptcode_.insert(ptcode_.end(), P.ptcode.begin(), P.ptcode.end());
    lsplit_[s] = std::move(P.lsplit);
    if (P.moms.size() % MOM_STRIDE) throw DeviceError("patch moments out of step");
    const uint64_t mom_base = moms_.size() / MOM_STRIDE;
    moms_.insert(moms_.end(), P.moms.begin(), P.moms.end());
    for (size_t g = 0; g < P.segs.size(); ++g) {
      SegDev sg = P.segs[g];
      sg.rec_off += rec_base[s];
      sg.set = static_cast<uint32_t>(s);
      if (mom_base + sg.patch_off > UINT32_MAX) throw DeviceError("too many patches");
      sg.mom_off = static_cast<uint32_t>(mom_base + sg.patch_off);
      uint32_t gid = static_cast<uint32_t>(segs_.size());
      if (P.seg_grid[g] < 7) {
        H.grids[P.seg_grid[g]].segs.push_back(gid);
        if (!P.seg_orig.empty()) H.grids[P.seg_grid[g]].orig.push_back(std::move(P.seg_orig[g]));
      }
      segs_.push_back(sg);
    }
    if (P.lane_e.size() != P.lane_phi.size()) throw DeviceError("lane earth vectors out of step");
    lane_phi_.insert(lane_phi_.end(), P.lane_phi.begin(), P.lane_phi.end());
    lane_e_.insert(lane_e_.end(), P.lane_e.begin(), P.lane_e.end());
    lane_meta_.insert(lane_meta_.end(), P.lane_meta.begin(), P.lane_meta.end());
    t_lane0_.insert(t_lane0_.end(), P.t_lane0.begin(), P.t_lane0.end());
    t_ndiv_.insert(t_ndiv_.end(), P.t_ndiv.begin(), P.t_ndiv.end());
    t_slfac_.insert(t_slfac_.end(), P.t_slfac.begin(), P.t_slfac.end());
    t_y_.insert(t_y_.end(), P.t_y.begin(), P.t_y.end());
    t_ferr_.insert(t_ferr_.end(), P.t_ferr.begin(), P.t_ferr.end());
    t_w_.insert(t_w_.end(), P.t_w.begin(), P.t_w.end());
    if (P.t_phase.size() != P.t_w.size()) throw DeviceError("time phases out of step");
    t_phase_.insert(t_phase_.end(), P.t_phase.begin(), P.t_phase.end());
    if (sd.flags & SET_ISCALE) any_iscale_ = true;
    if (!P.recs.empty())
      hip_check(hipMemcpy(d_recs + rec_base[s], P.recs.data(), P.recs.size(), hipMemcpyHostToDevice),
                "hipMemcpy records");
    packs[s].reset();  // release host staging as we go
  }

  // Host-side consistency checks of the per-set tables.
  for (size_t s = 0; s < ns; ++s) {
    const SetDev& sd = sets_[s];
    if (sd.lane_off % TILE || sd.n_lanes % TILE || sd.lane_off + sd.n_lanes > lane_phi_.size())
      throw DeviceError("set lanes out of range");
    if (sd.time_off + sd.n_times > t_lane0_.size()) throw DeviceError("set times out of range");
    if ((sd.flags & SET_VALID) && sd.special0 >= sd.n_lanes) throw DeviceError("special lanes out of range");
    for (uint32_t t = 0; t < sd.n_times; ++t) {
      const uint64_t end = static_cast<uint64_t>(t_lane0_[sd.time_off + t]) +
                           static_cast<uint64_t>(std::max(0, t_ndiv_[sd.time_off + t]));
      if ((sd.flags & SET_VALID) && end > sd.special0) throw DeviceError("time lanes out of range");
    }
  }

  // Tile class/grid-type masks.
  tile_cp_.assign(lane_meta_.size() / TILE, 0u);
  for (size_t l = 0; l < lane_meta_.size(); ++l) {
    uint32_t m = lane_meta_[l];
    if (m & META_VALID) tile_cp_[l / TILE] |= cp_bit(meta_class(m), meta_ptype(m));
  }

  d_.n_sets = static_cast<uint32_t>(ns);
  d_.max_set_lanes = 0;
  for (const SetDev& sd : sets_)
    if (sd.flags & SET_VALID) d_.max_set_lanes = std::max(d_.max_set_lanes, sd.special0);
  d_.recs = d_recs;
  upload();
  if (!mem_.arena) {
    hip_check(hipEventCreate(&ev_[0]), "hipEventCreate");
    hip_check(hipEventCreate(&ev_[1]), "hipEventCreate");
    hip_check(hipEventCreate(&ev_[2]), "hipEventCreate");
    hip_check(hipStreamCreateWithFlags(&side_.side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    hip_check(hipEventCreateWithFlags(&side_.fork, hipEventDisableTiming), "hipEventCreateWithFlags");
    hip_check(hipEventCreateWithFlags(&side_.join, hipEventDisableTiming), "hipEventCreateWithFlags");
  }
  build_items(ctx_.tuning.chunk);
}

Batch::~Batch() { release_all(); }

void Batch::release_all() {
  (void)hipSetDevice(ctx_.device());
  // Measured on ROCm 7.2: device memory freed after a graph launch with only
  // the launching stream synchronised stays allocated for the rest of the
  // process (about 230 MB of a 128-set batch's 270 MB, every batch); after a
  // device-wide synchronisation it is released. A batch in given memory
  // frees nothing and skips the wait.
  if (!mem_.arena) (void)hipDeviceSynchronize();
  release_graph();
  for (hipEvent_t& e : ev_) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (side_.side) (void)hipStreamSynchronize(side_.side);
  if (side_.fork) (void)hipEventDestroy(side_.fork);
  if (side_.join) (void)hipEventDestroy(side_.join);
  if (side_.side) (void)hipStreamDestroy(side_.side);
  side_ = SideStream{};
  item_free(d_items_);
  item_free(d_imoms_);
  item_free(d_partials_);
  item_free(d_limb_);
  item_free(const_cast<SetDev*>(d_.sets));
  item_free(const_cast<uint32_t*>(d_.chunk_info));
  d_items_ = nullptr;
  d_imoms_ = nullptr;
  d_partials_ = nullptr;
  d_limb_ = nullptr;
  d_.sets = nullptr;
  d_.chunk_info = nullptr;
  dfree_all();
}

void Batch::upload() {
  auto up = [&](auto& vec, auto** dst) {
    using T = typename std::remove_reference<decltype(vec)>::type::value_type;
    T* p = dalloc<T>(vec.size());
    if (!vec.empty())
      hip_check(hipMemcpy(p, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    *dst = p;
  };
  const SegDev* segs;
  up(segs_, const_cast<SegDev**>(&segs));
  d_.segs = segs;
  const double* lp;
  up(lane_phi_, const_cast<double**>(&lp));
  d_.lane_phi = lp;
  const float4* fp;
  up(lane_e_, const_cast<float4**>(&fp));
  d_.lane_e = fp;
  const uint32_t* u;
  up(lane_meta_, const_cast<uint32_t**>(&u));
  d_.lane_meta = u;
  up(t_lane0_, const_cast<uint32_t**>(&u));
  d_.t_lane0 = u;
  const int* ip;
  up(t_ndiv_, const_cast<int**>(&ip));
  d_.t_ndiv = ip;
  up(t_slfac_, const_cast<double**>(&lp));
  d_.t_slfac = lp;
  up(t_y_, const_cast<double**>(&lp));
  d_.t_y = lp;
  up(t_ferr_, const_cast<double**>(&lp));
  d_.t_ferr = lp;
  up(t_w_, const_cast<double**>(&lp));
  d_.t_w = lp;
  up(t_phase_, const_cast<double**>(&lp));
  d_.t_phase = lp;
  if (any_iscale_) {
    d_.lane_comp = dalloc<double>(lane_phi_.size() * 5);
    d_.comp = dalloc<double>(t_lane0_.size() * 5);
    d_.lsw = dalloc<double>(t_lane0_.size() * 6);
  }
  up(moms_, const_cast<double**>(&lp));
  d_.moms = lp;
  d_.lane_tot = dalloc<double>(lane_phi_.size());
  d_.flux = dalloc<double>(t_lane0_.size());
  d_.out = dalloc<SetOut>(sets_.size());
  // sets_ and chunk_info_ are (re)uploaded by build_items.
}

void Batch::release_graph() {
  if (graph_) {
    (void)hipGraphExecDestroy(graph_);
    graph_ = nullptr;
  }
  graph_scale_ = -1;
}

void Batch::build_items(uint32_t chunk) {
  // Items cover whole element patches, at most 64 of them.
  chunk = std::max<uint32_t>(PATCH, (chunk + PATCH - 1) / PATCH * PATCH);
  chunk = std::min<uint32_t>(chunk, MAX_ITEM_PATCHES * PATCH);
  if (mem_.arena && items_built_) throw DeviceError("the work items of a batch in given memory are built once");
  items_built_ = true;
  // Launches in flight still read the tables replaced below; the device-wide
  // wait is the one that lets the runtime release them (Batch::release_all).
  hip_check(mem_.arena ? hipStreamSynchronize(ctx_.stream()) : hipDeviceSynchronize(), "synchronize");
  release_graph();
  chunk_ = chunk;
  chunk_info_.clear();
  limb_chunks_.clear();
  limb_items_.clear();
  limb_idx_.clear();
  limb_lanes_.clear();
  lane_masks_.clear();
  pairs_limb_ = 0;
  limb_slots_ = 0;
  // Sequential pass: chunk tables and partial-row bases per set.
  struct SetPlan {
    size_t g0 = 0, g1 = 0;
    std::vector<uint32_t> chunk0;
    uint32_t nch = 0;
  };
  std::vector<SetPlan> plans(sets_.size());
  uint64_t prow = 0;
  size_t seg_i = 0;
  for (size_t s = 0; s < sets_.size(); ++s) {
    SetDev& sd = sets_[s];
    SetPlan& pl = plans[s];
    sd.chunk_off = static_cast<uint32_t>(chunk_info_.size());
    // Segments of this set are contiguous in segs_.
    pl.g0 = seg_i;
    while (seg_i < segs_.size() && segs_[seg_i].set == s) ++seg_i;
    pl.g1 = seg_i;
    pl.chunk0.resize(pl.g1 - pl.g0);
    for (size_t g = pl.g0; g < pl.g1; ++g) {
      pl.chunk0[g - pl.g0] = pl.nch;
      uint32_t c = (segs_[g].n + chunk - 1) / chunk;
      for (uint32_t k = 0; k < c; ++k) chunk_info_.push_back(segs_[g].use_mask | (segs_[g].kind << 16));
      pl.nch += c;
    }
    sd.n_chunks = pl.nch;
    sd.prow_base = prow;
    prow += static_cast<uint64_t>(sd.n_lanes / TILE) * pl.nch;
  }
  n_prows_ = prow;

  // Per-set items and limb tables, built in parallel (offsets set-relative
  // where they index the tables being built).
  struct SetItems {
    std::vector<ItemDev> items;
    std::vector<uint64_t> masks;
    std::vector<LimbChunk> lchunks;
    std::vector<LimbItem> litems;
    std::vector<uint16_t> lidx;
    std::vector<uint32_t> llanes;
    uint64_t pairs_limb = 0;
    uint64_t limb_slots = 0;
    std::string error;
  };
  std::vector<SetItems> out(sets_.size());
  const int limb = ctx_.tuning.limb;
  auto build_set = [&](size_t s) {
    const SetDev& sd = sets_[s];
    if (!(sd.flags & SET_VALID)) return;
    const SetPlan& pl = plans[s];
    SetItems& R = out[s];
    const uint32_t ntile = sd.n_lanes / TILE;
    const std::vector<SetPack::LaneSplit>& spl_all = lsplit_[s];
    // Scratch reused across chunks.
    std::vector<const SetPack::LaneSplit*> tbl;
    std::vector<std::vector<uint32_t>> plan(MAX_ITEM_PATCHES);
    std::vector<uint32_t> uni;
    std::vector<int32_t> uidx(sd.n_lanes, -1);
    struct Tent {
      ItemDev it;
      uint32_t m0, mn;  // masks of this item in `tmasks`
      bool limb_row;
    };
    std::vector<Tent> tent;
    std::vector<uint64_t> tmasks;
    for (size_t g = pl.g0; g < pl.g1; ++g) {
      const SegDev& sg = segs_[g];
      const uint32_t c = (sg.n + chunk - 1) / chunk;
      for (uint32_t k = 0; k < c; ++k) {
        const uint32_t e0 = k * chunk, ne = std::min(chunk, sg.n - k * chunk), np = ne / PATCH;
        const uint32_t p0 = sg.patch_off + e0 / PATCH;
        // Lane splits of this chunk's (patch, tile) pairs.
        tbl.assign(static_cast<size_t>(np) * ntile, nullptr);
        if (limb) {
          const uint32_t lo = p0 * ntile, hi = (p0 + np) * ntile;
          auto it0 = std::lower_bound(spl_all.begin(), spl_all.end(), lo,
                                      [](const SetPack::LaneSplit& a, uint32_t x) { return a.pt < x; });
          for (auto it = it0; it != spl_all.end() && it->pt < hi; ++it) tbl[it->pt - lo] = &*it;
        }
        for (uint32_t q = 0; q < np; ++q) plan[q].clear();
        tent.clear();
        tmasks.clear();
        for (uint32_t t = 0; t < ntile; ++t) {
          if (!(tile_cp_[sd.lane_off / TILE + t] & sg.use_mask)) continue;
          Tent tt{};
          ItemDev& it = tt.it;
          it.seg = static_cast<uint32_t>(g);
          it.e0 = e0;
          it.ne = ne;
          it.lane0 = sd.lane_off + t * TILE;
          it.prow = sd.prow_base + static_cast<uint64_t>(t) * pl.nch + pl.chunk0[g - pl.g0] + k;
          tt.m0 = static_cast<uint32_t>(tmasks.size());
          for (uint32_t q = 0; q < np; ++q) {
            const uint64_t ci = pt_base_[s] + static_cast<uint64_t>(p0 + q) * pt_tiles_[s] + t;
            if (ci >= ptcode_.size()) {
              R.error = "patch code out of range";
              return;
            }
            const uint8_t code = ptcode_[ci];
            if (code == PT_FREE) {
              const SetPack::LaneSplit* sp = tbl[static_cast<size_t>(q) * ntile + t];
              if (sp && __builtin_popcountll(sp->some) <= limb) {
                it.pmom |= 1ull << q;
                tmasks.push_back(sp->all);
                for (uint64_t m = sp->some; m; m &= m - 1ull)
                  plan[q].push_back(it.lane0 + static_cast<uint32_t>(__builtin_ctzll(m)));
                if (sp->some) tt.limb_row = true;
                continue;
              }
            }
            if (code == PT_FULL || code == PT_FREE) it.todo |= 1ull << q;
            if (code == PT_FREE) it.free |= 1ull << q;
            if (code == PT_MOM) it.mom |= 1ull << q;
          }
          tt.mn = static_cast<uint32_t>(tmasks.size()) - tt.m0;
          tent.push_back(tt);
        }
        // The chunk's limb lanes; too many for one workgroup's LDS: the
        // patches go back to the items' element loops.
        uni.clear();
        for (uint32_t q = 0; q < np; ++q) uni.insert(uni.end(), plan[q].begin(), plan[q].end());
        std::sort(uni.begin(), uni.end());
        uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
        if (uni.size() > LIMB_UNION_MAX) {
          for (Tent& tt : tent) {
            tt.it.todo |= tt.it.pmom;
            tt.it.free |= tt.it.pmom;
            tt.it.pmom = 0;
            tt.mn = 0;
            tt.limb_row = false;
          }
          uni.clear();
        }
        if (!uni.empty()) {
          LimbChunk lc{};
          lc.seg = static_cast<uint32_t>(g);
          lc.chunk = pl.chunk0[g - pl.g0] + k;
          // Fixed-point scale: a lane's total over the chunk's limb
          // patches is at most sum |w| * max |I(m)|, m in [0, 1].
          double imax = 1.0;
          if (sg.kind <= KIND_STAR2) {
            const float* a = sg.kind == KIND_STAR1 ? sd.ldc1 : sd.ldc2;
            imax += std::fabs(a[0]) + std::fabs(a[1]) + std::fabs(a[2]) + std::fabs(a[3]);
          } else if (sg.kind != KIND_SPOT) {
            imax += std::fabs(sd.lin_disc) + std::fabs(sd.quad_disc);
          }
          double wsum = 0.0;
          for (uint32_t q = 0; q < np; ++q) {
            if (plan[q].empty()) continue;
            const size_t mi = static_cast<size_t>(sg.mom_off + e0 / PATCH + q) * MOM_STRIDE + 19;
            if (mi >= moms_.size()) {
              R.error = "limb patch moments out of range";
              return;
            }
            wsum += moms_[mi];
          }
          const double bound = 1.01 * imax * wsum;
          int ex = 0;
          if (bound > 0.0) (void)std::frexp(bound, &ex);  // bound < 2^ex
          lc.shift = bound > 0.0 ? 61 - ex : 0;
          lc.item0 = static_cast<uint32_t>(R.litems.size());
          lc.lane_off = static_cast<uint32_t>(R.llanes.size());
          lc.n_lanes = static_cast<uint32_t>(uni.size());
          for (size_t u = 0; u < uni.size(); ++u) uidx[uni[u] - sd.lane_off] = static_cast<int32_t>(u);
          R.llanes.insert(R.llanes.end(), uni.begin(), uni.end());
          for (uint32_t q = 0; q < np; ++q) {
            const std::vector<uint32_t>& v = plan[q];
            for (size_t i0 = 0; i0 < v.size();) {
              // Up to 64 lanes per item. A remainder of at most 32 (16)
              // lanes takes a split item: 2 (4) lane groups holding the
              // same lanes, each summing a half (quarter) of the patch's
              // elements.
              const size_t rem = v.size() - i0;
              const uint32_t split = rem <= TILE / 4 ? 2u : (rem <= TILE / 2 ? 1u : 0u);
              const size_t gl = TILE >> split;
              R.litems.push_back(LimbItem{e0 + q * PATCH, split});
              for (uint32_t g = 0; g < (1u << split); ++g)
                for (size_t i = i0; i < i0 + gl; ++i) {
                  uint16_t u = 0xFFFFu;
                  if (i < v.size()) {
                    u = static_cast<uint16_t>(uidx[v[i] - sd.lane_off]);
                    if (g == 0) R.pairs_limb += PATCH;
                  }
                  R.lidx.push_back(u);
                }
              R.limb_slots += static_cast<uint64_t>(TILE) * (PATCH >> split);
              i0 += gl;
            }
          }
          lc.n_items = static_cast<uint32_t>(R.litems.size()) - lc.item0;
          R.lchunks.push_back(lc);
        }
        for (Tent& tt : tent) {
          // An item adding zero to every lane is not launched: its partial
          // row stays at the zero written once below. A row that limb
          // lanes add to is rewritten by its item on every launch.
          if (tt.it.todo == 0 && tt.it.mom == 0 && tt.it.pmom == 0 && !tt.limb_row) continue;
          tt.it.lm_off = static_cast<uint32_t>(R.masks.size());
          R.masks.insert(R.masks.end(), tmasks.begin() + tt.m0, tmasks.begin() + tt.m0 + tt.mn);
          R.items.push_back(tt.it);
        }
      }
    }
  };
  {
    int nthreads = static_cast<int>(std::thread::hardware_concurrency());
    if (const char* env = std::getenv("LCURVE_HIP_THREADS")) nthreads = std::max(1, std::atoi(env));
    nthreads = std::max(1, std::min<int>(nthreads, static_cast<int>(sets_.size())));
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
      pool.emplace_back([&] {
        for (size_t s; (s = next.fetch_add(1)) < sets_.size();) build_set(s);
      });
    for (auto& th : pool) th.join();
  }
  // Concatenate, moving the set-relative offsets.
  std::vector<ItemDev> items;
  {
    size_t ni = 0, nm = 0, nc = 0, nli = 0, nll = 0;
    for (const SetItems& R : out) {
      if (!R.error.empty()) throw DeviceError(R.error);
      ni += R.items.size();
      nm += R.masks.size();
      nc += R.lchunks.size();
      nli += R.litems.size();
      nll += R.llanes.size();
    }
    if (nm > UINT32_MAX || nli > UINT32_MAX / TILE || nll > UINT32_MAX) throw DeviceError("limb tables too large");
    items.reserve(ni);
    lane_masks_.reserve(nm);
    limb_chunks_.reserve(nc);
    limb_items_.reserve(nli);
    limb_idx_.reserve(nli * TILE);
    limb_lanes_.reserve(nll);
    for (SetItems& R : out) {
      const uint32_t mbase = static_cast<uint32_t>(lane_masks_.size());
      const uint32_t ibase = static_cast<uint32_t>(limb_items_.size());
      const uint32_t lbase = static_cast<uint32_t>(limb_lanes_.size());
      for (ItemDev& it : R.items) {
        it.lm_off += mbase;
        items.push_back(it);
      }
      for (LimbChunk& lc : R.lchunks) {
        lc.item0 += ibase;
        lc.lane_off += lbase;
        limb_chunks_.push_back(lc);
      }
      lane_masks_.insert(lane_masks_.end(), R.masks.begin(), R.masks.end());
      limb_items_.insert(limb_items_.end(), R.litems.begin(), R.litems.end());
      limb_idx_.insert(limb_idx_.end(), R.lidx.begin(), R.lidx.end());
      limb_lanes_.insert(limb_lanes_.end(), R.llanes.begin(), R.llanes.end());
      pairs_limb_ += R.pairs_limb;
      limb_slots_ += R.limb_slots;
      R = SetItems();
    }
  }
  // Items of beamed / lensed star segments go last: the kernel for the
  // others is compiled without those paths.
  auto needs_fx = [&](const ItemDev& it) {
    const SegDev& sg = segs_[it.seg];
    return sg.kind <= KIND_STAR2 && (sg.flags & SEG_POSN) != 0;
  };
  const auto mid = std::stable_partition(items.begin(), items.end(), [&](const ItemDev& it) { return !needs_fx(it); });
  // Of the others, items with an eclipse-tested element-loop patch of a
  // segment with 1..LDS_ECL_MAX eclipse intervals go second: their kernel
  // stages those intervals in LDS and needs more registers than the first
  // one. (Items whose patches all skip the eclipse test stage records
  // without intervals in the first kernel.)
  auto staged_ecl = [&](const ItemDev& it) {
    const SegDev& sg = segs_[it.seg];
    return (it.todo & ~it.free) != 0 && sg.necl >= 1 && sg.necl <= LDS_ECL_MAX;
  };
  const auto mid2 = std::stable_partition(items.begin(), mid, [&](const ItemDev& it) { return !staged_ecl(it); });
  n_items_plain_ = static_cast<uint64_t>(mid2 - items.begin());
  n_items_ecl_ = static_cast<uint64_t>(mid - mid2);
  n_items_ = items.size();
  // Item moments: the sum, in patch order, of the moments of an item's
  // PT_MOM patches, so that the kernel evaluates one moment sum per item and
  // lane for them; then, for its split patches (pmom), one sum per distinct
  // lane mask (in order of first appearance), whose masks replace the
  // per-patch ones in place. FP64 on the host, across the setup threads.
  {
    uint64_t n_im = 0;
    for (ItemDev& it : items) {
      uint32_t g = 0;
      if (it.pmom) {
        const uint32_t np = static_cast<uint32_t>(__builtin_popcountll(it.pmom));
        if (static_cast<uint64_t>(it.lm_off) + np > lane_masks_.size()) throw DeviceError("item lane masks out of range");
        uint64_t seen[MAX_ITEM_PATCHES];
        for (uint32_t i = 0; i < np; ++i) {
          const uint64_t m = lane_masks_[it.lm_off + i];
          uint32_t k = 0;
          while (k < g && seen[k] != m) ++k;
          if (k == g) seen[g++] = m;
        }
      }
      it.ngrp = g;
      it.pad = 0;
      const uint64_t ne = (it.mom ? 1u : 0u) + g;
      it.imom = ne ? static_cast<uint32_t>(n_im) : 0u;
      n_im += ne;
      if (n_im > UINT32_MAX) throw DeviceError("too many item moments");
    }
    item_moms_.assign(std::max<uint64_t>(n_im, 1) * MOM_STRIDE, 0.0);
    for (const ItemDev& it : items)
      if ((it.mom || it.pmom) &&
          (it.seg >= segs_.size() ||
           static_cast<uint64_t>(segs_[it.seg].mom_off) + (it.e0 + it.ne) / PATCH > moms_.size() / MOM_STRIDE))
        throw DeviceError("item moments out of range");
    int nthreads = static_cast<int>(std::thread::hardware_concurrency());
    if (const char* env = std::getenv("LCURVE_HIP_THREADS")) nthreads = std::max(1, std::atoi(env));
    nthreads = std::max(1, nthreads);
    const size_t blk = 4096;
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
      pool.emplace_back([&] {
        for (size_t b; (b = next.fetch_add(blk)) < items.size();)
          for (size_t i = b; i < std::min(items.size(), b + blk); ++i) {
            const ItemDev& it = items[i];
            if (!it.mom && !it.pmom) continue;
            double* o = item_moms_.data() + static_cast<size_t>(it.imom) * MOM_STRIDE;
            const double* m0 = moms_.data() + (static_cast<size_t>(segs_[it.seg].mom_off) + it.e0 / PATCH) * MOM_STRIDE;
            auto add = [&](double* dst, uint32_t q) {
              const double* a = m0 + static_cast<size_t>(q) * MOM_STRIDE;
              for (uint32_t c = 0; c < MOM_STRIDE; ++c) dst[c] += a[c];
            };
            if (it.mom) {
              for (uint64_t mm = it.mom; mm; mm &= mm - 1ull) add(o, static_cast<uint32_t>(__builtin_ctzll(mm)));
              o += MOM_STRIDE;
            }
            if (it.pmom) {
              // This item's masks, one per pmom patch in bit order; only
              // this thread touches them.
              uint64_t* mk = lane_masks_.data() + it.lm_off;
              uint64_t gm[MAX_ITEM_PATCHES];
              uint32_t g = 0, r = 0;
              for (uint64_t mm = it.pmom; mm; mm &= mm - 1ull, ++r) {
                const uint64_t m = mk[r];
                uint32_t k = 0;
                while (k < g && gm[k] != m) ++k;
                if (k == g) gm[g++] = m;
                add(o + static_cast<size_t>(k) * MOM_STRIDE, static_cast<uint32_t>(__builtin_ctzll(mm)));
              }
              for (uint32_t k = 0; k < g; ++k) mk[k] = gm[k];
            }
          }
      });
    for (auto& th : pool) th.join();
  }
  pairs_eval_ = pairs_mom_ = 0;
  mom_evals_ = 0;
  for (const ItemDev& it : items) {
    pairs_eval_ += static_cast<uint64_t>(__builtin_popcountll(it.todo)) * PATCH * TILE;
    pairs_mom_ += static_cast<uint64_t>(__builtin_popcountll(it.mom) + __builtin_popcountll(it.pmom)) * PATCH * TILE;
    // Moment evaluations as the kernel does them: one for the item's PT_MOM
    // patches, one per lane-mask group of its split patches.
    mom_evals_ += ((it.mom ? 1u : 0u) + static_cast<uint64_t>(it.ngrp)) * TILE;
  }
  pairs_eval_ += pairs_limb_;
  // Host-side bounds check of every item against the buffers it indexes.
  for (const ItemDev& it : items) {
    if (it.seg >= segs_.size()) throw DeviceError("item segment out of range");
    const SegDev& sg = segs_[it.seg];
    const SetDev& sd = sets_[sg.set];
    if (static_cast<uint64_t>(it.e0) + it.ne > sg.n || it.ne == 0) throw DeviceError("item elements out of range");
    if (it.lane0 % TILE || it.lane0 < sd.lane_off || it.lane0 + TILE > sd.lane_off + sd.n_lanes ||
        it.lane0 + TILE > lane_phi_.size())
      throw DeviceError("item lanes out of range");
    if (it.prow >= n_prows_ || it.prow < sd.prow_base) throw DeviceError("item partial row out of range");
    const uint64_t stride = rec_stride(sg.necl, (sg.flags & SEG_POSN) != 0);
    if (sg.rec_off + (static_cast<uint64_t>(it.e0) + it.ne) * stride > rec_bytes_)
      throw DeviceError("item records out of range");
    if (sg.necl > static_cast<uint32_t>(MAX_ECL)) throw DeviceError("segment eclipse count out of range");
    if (it.e0 % PATCH || it.ne % PATCH || it.ne / PATCH > MAX_ITEM_PATCHES)
      throw DeviceError("item not aligned to patches");
    const uint64_t live = it.ne / PATCH == 64 ? ~0ull : (1ull << (it.ne / PATCH)) - 1ull;
    if ((it.todo & ~live) || (it.free & ~it.todo) || (it.mom & ~live) || (it.mom & it.todo))
      throw DeviceError("item patch masks out of range");
    if ((it.mom || it.pmom) && static_cast<uint64_t>(sg.mom_off) + (it.e0 + it.ne) / PATCH > moms_.size() / MOM_STRIDE)
      throw DeviceError("item moments out of range");
    if ((it.mom || it.pmom) &&
        (static_cast<uint64_t>(it.imom) + (it.mom ? 1u : 0u) + it.ngrp) * MOM_STRIDE > item_moms_.size())
      throw DeviceError("item moment entry out of range");
    if ((it.pmom == 0) != (it.ngrp == 0) || it.ngrp > static_cast<uint32_t>(__builtin_popcountll(it.pmom)))
      throw DeviceError("item moment groups out of range");
    if ((it.pmom & ~live) || (it.pmom & (it.todo | it.mom)) ||
        static_cast<uint64_t>(it.lm_off) + __builtin_popcountll(it.pmom) > lane_masks_.size())
      throw DeviceError("item lane masks out of range");
  }
  for (const LimbChunk& lc : limb_chunks_) {
    if (lc.seg >= segs_.size()) throw DeviceError("limb chunk segment out of range");
    const SegDev& sg = segs_[lc.seg];
    const SetDev& sd = sets_[sg.set];
    if (!(sd.flags & SET_VALID) || lc.chunk >= sd.n_chunks) throw DeviceError("limb chunk out of range");
    if ((sg.flags & SEG_POSN) || sg.necl > static_cast<uint32_t>(MAX_ECL)) throw DeviceError("limb chunk segment kind");
    if (lc.n_lanes == 0 || lc.n_lanes > LIMB_UNION_MAX ||
        static_cast<uint64_t>(lc.lane_off) + lc.n_lanes > limb_lanes_.size())
      throw DeviceError("limb chunk lanes out of range");
    if (static_cast<uint64_t>(lc.item0) + lc.n_items > limb_items_.size()) throw DeviceError("limb items out of range");
    if (lc.shift < -1000 || lc.shift > 1000) throw DeviceError("limb fixed-point scale out of range");
    for (uint32_t u = 0; u < lc.n_lanes; ++u) {
      const uint32_t gl = limb_lanes_[lc.lane_off + u];
      if (gl < sd.lane_off || gl >= sd.lane_off + sd.n_lanes) throw DeviceError("limb lane out of range");
      const uint64_t row = sd.prow_base + static_cast<uint64_t>((gl - sd.lane_off) / TILE) * sd.n_chunks + lc.chunk;
      if (row >= n_prows_) throw DeviceError("limb partial row out of range");
    }
    const uint64_t stride = rec_stride(sg.necl, false);
    for (uint32_t i = lc.item0; i < lc.item0 + lc.n_items; ++i) {
      const LimbItem& li = limb_items_[i];
      if (li.split > 2) throw DeviceError("limb item split out of range");
      if (li.e0 % PATCH || static_cast<uint64_t>(li.e0) + PATCH > sg.n ||
          sg.rec_off + (static_cast<uint64_t>(li.e0) + PATCH) * stride > rec_bytes_)
        throw DeviceError("limb item records out of range");
      for (uint32_t j = 0; j < TILE; ++j) {
        const uint16_t u = limb_idx_[static_cast<size_t>(i) * TILE + j];
        if (u != 0xFFFFu && u >= lc.n_lanes) throw DeviceError("limb item lane out of range");
      }
    }
  }
  if (limb_idx_.size() != limb_items_.size() * TILE) throw DeviceError("limb index table out of step");
  if (n_items_ > 0xFFFFFFFFull / TILE) throw DeviceError("batch too large for one launch");

  item_free(d_items_);
  item_free(d_imoms_);
  item_free(d_partials_);
  item_free(d_limb_);
  d_items_ = nullptr;
  d_imoms_ = nullptr;
  d_partials_ = nullptr;
  d_limb_ = nullptr;
  d_items_ = static_cast<ItemDev*>(item_alloc(std::max<size_t>(items.size(), 1) * sizeof(ItemDev), "items"));
  d_imoms_ = static_cast<double*>(item_alloc(item_moms_.size() * sizeof(double), "item moments"));
  hip_check(hipMemcpy(d_imoms_, item_moms_.data(), item_moms_.size() * sizeof(double), hipMemcpyHostToDevice),
            "hipMemcpy item moments");
  d_.imoms = d_imoms_;
  // Limb tables and item lane masks, one allocation.
  {
    const size_t b_ch = align256(limb_chunks_.size() * sizeof(LimbChunk));
    const size_t b_it = align256(limb_items_.size() * sizeof(LimbItem));
    const size_t b_ix = align256(limb_idx_.size() * sizeof(uint16_t));
    const size_t b_ln = align256(limb_lanes_.size() * sizeof(uint32_t));
    const size_t b_lm = align256(lane_masks_.size() * sizeof(uint64_t));
    const size_t total = std::max<size_t>(b_ch + b_it + b_ix + b_ln + b_lm, 256);
    d_limb_ = item_alloc(total, "limb tables");
    char* base = static_cast<char*>(d_limb_);
    auto put = [&](const void* src, size_t bytes, size_t at) {
      if (bytes) hip_check(hipMemcpy(base + at, src, bytes, hipMemcpyHostToDevice), "hipMemcpy limb tables");
    };
    put(limb_chunks_.data(), limb_chunks_.size() * sizeof(LimbChunk), 0);
    put(limb_items_.data(), limb_items_.size() * sizeof(LimbItem), b_ch);
    put(limb_idx_.data(), limb_idx_.size() * sizeof(uint16_t), b_ch + b_it);
    put(limb_lanes_.data(), limb_lanes_.size() * sizeof(uint32_t), b_ch + b_it + b_ix);
    put(lane_masks_.data(), lane_masks_.size() * sizeof(uint64_t), b_ch + b_it + b_ix + b_ln);
    d_.n_limb_chunks = static_cast<uint32_t>(limb_chunks_.size());
    d_.limb_chunks = reinterpret_cast<const LimbChunk*>(base);
    d_.limb_items = reinterpret_cast<const LimbItem*>(base + b_ch);
    d_.limb_idx = reinterpret_cast<const uint16_t*>(base + b_ch + b_it);
    d_.limb_lanes = reinterpret_cast<const uint32_t*>(base + b_ch + b_it + b_ix);
    d_.lane_masks = reinterpret_cast<const uint64_t*>(base + b_ch + b_it + b_ix + b_ln);
  }
  d_partials_ = static_cast<double*>(item_alloc(std::max<uint64_t>(n_prows_, 1) * TILE * sizeof(double), "partials"));
  // Rows of items that are not launched must read as zero; launched items
  // rewrite theirs on every launch. Cleared here, outside any graph (given
  // memory keeps whatever an earlier batch left in it).
  hip_check(hipMemset(d_partials_, 0, std::max<uint64_t>(n_prows_, 1) * TILE * sizeof(double)), "hipMemset partials");
  if (!items.empty())
    hip_check(hipMemcpy(d_items_, items.data(), items.size() * sizeof(ItemDev), hipMemcpyHostToDevice),
              "hipMemcpy items");
  // Per-set table and chunk info (reallocated: sizes change with chunk).
  SetDev* dsets = static_cast<SetDev*>(item_alloc(std::max<size_t>(sets_.size(), 1) * sizeof(SetDev), "set table"));
  uint32_t* dci = static_cast<uint32_t*>(item_alloc(std::max<size_t>(chunk_info_.size(), 1) * sizeof(uint32_t), "chunk table"));
  if (!sets_.empty())
    hip_check(hipMemcpy(dsets, sets_.data(), sets_.size() * sizeof(SetDev), hipMemcpyHostToDevice), "hipMemcpy");
  if (!chunk_info_.empty())
    hip_check(hipMemcpy(dci, chunk_info_.data(), chunk_info_.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
              "hipMemcpy");
  item_free(const_cast<SetDev*>(d_.sets));
  item_free(const_cast<uint32_t*>(d_.chunk_info));
  d_.sets = dsets;
  d_.chunk_info = dci;
  d_.items = d_items_;
  d_.n_items = static_cast<uint32_t>(n_items_);
  d_.n_items_plain = static_cast<uint32_t>(n_items_plain_);
  d_.n_items_ecl = static_cast<uint32_t>(n_items_ecl_);
  d_.partials = d_partials_;
}

uint64_t Batch::visits() const {
  uint64_t v = 0;
  for (const SetHost& h : host_) v += h.visits;
  return v;
}

Batch::Tables Batch::tables() const {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hip_check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
  Tables t;
  t.recs.resize(rec_bytes_);
  if (rec_bytes_)
    hip_check(hipMemcpy(t.recs.data(), d_.recs, rec_bytes_, hipMemcpyDeviceToHost), "hipMemcpy records");
  t.segs.resize(segs_.size());
  if (!segs_.empty())
    hip_check(hipMemcpy(t.segs.data(), d_.segs, segs_.size() * sizeof(SegDev), hipMemcpyDeviceToHost),
              "hipMemcpy segments");
  t.moms = moms_;
  t.imoms = item_moms_;
  t.codes = ptcode_;
  t.lane_masks = lane_masks_;
  t.limb_chunks = limb_chunks_;
  t.limb_items = limb_items_;
  t.limb_idx = limb_idx_;
  t.limb_lanes = limb_lanes_;
  t.items.resize(n_items_);
  if (n_items_)
    hip_check(hipMemcpy(t.items.data(), d_items_, n_items_ * sizeof(ItemDev), hipMemcpyDeviceToHost),
              "hipMemcpy items");
  return t;
}

void Batch::launch(int scale) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hipStream_t st = ctx_.stream();
  if (use_graph) {
    if (!graph_ || graph_scale_ != scale) {
      release_graph();
      hipGraph_t g;
      hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      hipError_t e1 = launch_flux_items(d_, st, &side_);
      hipError_t e2 = launch_assemble(d_, scale, st);
      hipError_t e3 = hipStreamEndCapture(st, &g);
      hip_check(e1, "capture flux kernel");
      hip_check(e2, "capture assemble kernel");
      hip_check(e3, "hipStreamEndCapture");
      hipError_t e4 = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      hip_check(e4, "hipGraphInstantiate");
      graph_scale_ = scale;
    }
    hip_check(hipGraphLaunch(graph_, st), "hipGraphLaunch");
  } else {
    hip_check(launch_flux_items(d_, st, &side_), "flux kernel launch");
    hip_check(launch_assemble(d_, scale, st), "assemble kernel launch");
  }
}

void Batch::run(int scale, double* chisq, double* wnok, double* sfac, double* wdwarf, double* flux) {
  launch(scale);
  hip_check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
  std::vector<SetOut>& out = last_out_;
  out.resize(sets_.size());
  if (!out.empty())
    hip_check(hipMemcpy(out.data(), d_.out, out.size() * sizeof(SetOut), hipMemcpyDeviceToHost), "hipMemcpy out");
  for (size_t s = 0; s < out.size(); ++s) {
    if (chisq) chisq[s] = out[s].chisq;
    if (wnok) wnok[s] = out[s].wnok;
    if (sfac) sfac[s] = out[s].scale;
    if (wdwarf) wdwarf[s] = out[s].wdwarf;
  }
  if (flux && !t_lane0_.empty())
    hip_check(hipMemcpy(flux, d_.flux, t_lane0_.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy flux");
}

void Batch::time(int scale, int reps, double* ms_flux, double* ms_total) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hipStream_t st = ctx_.stream();
  double tf = 0.0;
  hip_check(hipEventRecord(ev_[2], st), "hipEventRecord");
  for (int r = 0; r < reps; ++r) {
    hip_check(hipEventRecord(ev_[0], st), "hipEventRecord");
    hip_check(launch_flux_items(d_, st, &side_), "flux kernel launch");
    hip_check(hipEventRecord(ev_[1], st), "hipEventRecord");
    hip_check(launch_assemble(d_, scale, st), "assemble kernel launch");
    hip_check(hipEventSynchronize(ev_[1]), "hipEventSynchronize");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
    tf += ms;
  }
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  hipEvent_t end;
  hip_check(hipEventCreate(&end), "hipEventCreate");
  hip_check(hipEventRecord(end, st), "hipEventRecord");
  hip_check(hipEventSynchronize(end), "hipEventSynchronize");
  float total = 0.0f;
  hip_check(hipEventElapsedTime(&total, ev_[2], end), "hipEventElapsedTime");
  (void)hipEventDestroy(end);
  if (ms_flux) *ms_flux = reps > 0 ? tf / reps : 0.0;
  if (ms_total) *ms_total = reps > 0 ? total / reps : 0.0;
}

void Batch::visibility(int set, int grid, const double* phases, int nphase, uint8_t* mask,
                       uint32_t* n_out) {
  if (set < 0 || static_cast<size_t>(set) >= host_.size()) throw std::invalid_argument("set out of range");
  if (grid < 0 || grid > 6) throw std::invalid_argument("grid out of range");
  // An aliased coarse grid is the fine grid (the reference clones it).
  if ((grid == 1 && host_[set].star1_alias) || (grid == 3 && host_[set].star2_alias)) grid -= 1;
  const SetHost::GridMap& gm = host_[set].grids[grid];
  *n_out = gm.n;
  if (!mask || nphase <= 0 || gm.n == 0) return;
  if (gm.orig.size() != gm.segs.size()) throw std::invalid_argument("element maps were not kept for this batch");
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hip_check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
  double* dph = nullptr;
  hip_check(hipMalloc(reinterpret_cast<void**>(&dph), nphase * sizeof(double)), "hipMalloc");
  hip_check(hipMemcpy(dph, phases, nphase * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  for (size_t k = 0; k < gm.segs.size(); ++k) {
    const SegDev& sg = segs_[gm.segs[k]];
    size_t bytes = static_cast<size_t>(sg.n) * nphase;
    uint8_t* dm = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void**>(&dm), std::max<size_t>(bytes, 1)), "hipMalloc");
    hip_check(launch_visibility(d_, sg, dph, static_cast<uint32_t>(nphase), dm, ctx_.stream()), "visibility launch");
    std::vector<uint8_t> hm(bytes);
    hip_check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
    hip_check(hipMemcpy(hm.data(), dm, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    (void)hipFree(dm);
    const std::vector<uint32_t>& orig = gm.orig[k];
    for (int p = 0; p < nphase; ++p)
      for (uint32_t e = 0; e < sg.n; ++e)
        if (orig[e] != UINT32_MAX)  // padding records have no grid element
          mask[static_cast<size_t>(p) * gm.n + orig[e]] = hm[static_cast<size_t>(p) * sg.n + e];
  }
  (void)hipFree(dph);
}

void Batch::set_obs(const double* flux2d, const double* ferr2d, const double* weight2d) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hip_check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
  const size_t n = t_y_.size();
  if (n != sets_.size() * n_times_) throw std::invalid_argument("time tables do not match the set count");
  auto put = [&](const double* src, std::vector<double>& host, const double* dev) {
    if (!src) return;
    host.assign(src, src + n);
    hip_check(hipMemcpy(const_cast<double*>(dev), host.data(), n * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy observations");
  };
  put(flux2d, t_y_, d_.t_y);
  put(ferr2d, t_ferr_, d_.t_ferr);
  put(weight2d, t_w_, d_.t_w);
}

void Batch::set_curve_terms(const double* slfac, double third) {
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hip_check(hipStreamSynchronize(ctx_.stream()), "hipStreamSynchronize");
  const size_t n = t_slfac_.size();
  if (n != sets_.size() * n_times_) throw std::invalid_argument("time tables do not match the set count");
  if (slfac) {
    t_slfac_.assign(slfac, slfac + n);
    hip_check(hipMemcpy(const_cast<double*>(d_.t_slfac), t_slfac_.data(), n * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy polynomial factors");
  }
  for (SetDev& sd : sets_) sd.third = third;
  if (!sets_.empty())
    hip_check(hipMemcpy(const_cast<SetDev*>(d_.sets), sets_.data(), sets_.size() * sizeof(SetDev), hipMemcpyHostToDevice),
              "hipMemcpy set table");
}

void Batch::autotune(int scale) {
  if (mem_.arena) throw DeviceError("a batch in given memory is not autotuned");
  // Limb lanes at the context's threshold, or off (a small batch may not
  // repay the extra launch).
  struct Cand { uint32_t chunk; bool limb; };
  const Cand cands[] = {{256, true},  {512, true},  {1024, true},  {2048, true},  {4096, true},
                        {256, false}, {512, false}, {1024, false}, {2048, false}, {4096, false}};
  const int limb = ctx_.tuning.limb;
  Tuning best = ctx_.tuning;
  double best_ms = 1e300;
  for (const Cand& c : cands) {
    if (!c.limb && limb == 0) continue;  // already covered
    ctx_.tuning.chunk = c.chunk;
    ctx_.tuning.limb = c.limb ? limb : 0;
    build_items(c.chunk);
    double ms_f = 0, ms_t = 0;
    time(scale, 1, &ms_f, &ms_t);  // warm
    time(scale, 3, &ms_f, &ms_t);
    // The whole pass counts: smaller chunks can speed the flux kernels but
    // give the assembly more partial rows to sum.
    if (ms_t < best_ms) {
      best_ms = ms_t;
      best = ctx_.tuning;
    }
  }
  best.tuned = true;
  ctx_.tuning = best;
  build_items(best.chunk);
  // The limb threshold stays the caller's for batches built later.
  ctx_.tuning.limb = limb;
}

}  // namespace lch
