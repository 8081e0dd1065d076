// msl_serve.cpp -- serving engine: a mix of independent multislice jobs of
// uneven grid size, slice count, method and probe count.
//
// Jobs are grouped into classes that share every k-space table (grid size;
// for finite differences also the step size, which differs between plane-wave
// and probe runs, U.cu:2633-2634).  A class runs in batches cut to a device
// memory budget.  In a batch the waves of all its jobs sit in one array,
// ordered by slice count (longest first), so the waves still running at slice
// i are a prefix of the array: every slice is one row launch and one column
// launch over that prefix, with each wave reading the transmission of its own
// job through a per-wave slot index.  A job's potentials are synthesised per
// slice from its own binned atoms, as in MultisliceStepConv (U.cu:2085-2124)
// or MultisliceStep (U.cu:2040-2083).
//
// Grouped mode (msl_serve_params.group_sizes): conventional jobs of every
// grid size 256..2048 share their batches; per slice one potential launch
// covers every running job and one row and one column launch every running
// wave, whatever its size (one segment per size, k_pass.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "msl_host.hpp"

using namespace msl;
using namespace mslh;

struct msl_server {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SimFree {
  void operator()(msl_sim* s) const { msl_destroy(s); }
};
using SimPtr = std::unique_ptr<msl_sim, SimFree>;

struct Job {
  const msl_job* d = nullptr;
  SimPtr s;       // the job's handle, on the server's stream
  bool stem = false;
  int slot = -1;  // transmission / potential slot inside the running batch
};

// a job's waves [p0, p0 + cnt) (a plane-wave job has one wave)
struct Item {
  int job, p0, cnt, nslices;
};
// an item of a grouped batch and its size class
struct GItem {
  int cls;
  Item it;
};

struct Stats {
  double setup = 0, prop = 0, batches = 0, waves = 0, wave_slices = 0, launches = 0, peak = 0;
};

class Runner {
 public:
  Runner(msl_server* srv, const msl_serve_params& p, std::vector<Job>& jobs, Stats& st)
      : srv_(srv), p_(p), jobs_(jobs), st_(st) {}

  // switch the per-size state to the tables of cs
  void use_class(msl_sim* cs, bool fd) {
    cs_ = cs;
    fd_ = fd;
    n_ = cs->resX;
    np_ = (long long)n_ * n_;
    ntiles_ = n_ / col_tile(n_);
    prune_ = !fd && cs->prune_tiles > 0;
    fd_prune_ = fd && cs->prune_tiles > 0;
    for (int x = n_ / 4; x < 3 * n_ / 4 && fd_prune_; ++x)
      if (std::sqrt(cs->hkx[x] * cs->hkx[x] + 0.0f) <= cs->kmax) fd_prune_ = false;
  }

  // jobs of one class, sharing the tables of cs
  void run_class(const std::vector<int>& ids, msl_sim* cs, bool fd) {
    use_class(cs, fd);
    ensure_prop(cs, fd ? cs->dzFD : p_.dz);
    bool probes = false;
    for (int j : ids) probes = probes || jobs_[j].d->nprobe > 0;
    if (probes) probe_norm();
    else scratch_.ensure(np_);
    const long long budget = p_.budget_bytes > 0 ? p_.budget_bytes : (32LL << 30);
    const long long wave_b = np_ * 8 * (fd ? (fd_prune_ ? 6 : 5) : 1), slot_b = np_ * 8 * (fd ? 2 : 1);
    const long long maxw = std::max<long long>(1, (budget - slot_b) / (wave_b + slot_b));
    std::vector<Item> items;
    for (int j : ids) {
      const msl_job* d = jobs_[j].d;
      if (d->nprobe <= 0) {
        items.push_back({j, 0, 1, d->nslices});
      } else {
        for (int p0 = 0; p0 < d->nprobe; p0 += (int)maxw)
          items.push_back({j, p0, (int)std::min<long long>(maxw, d->nprobe - p0), d->nslices});
      }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.nslices > b.nslices; });
    // greedy batches in that order (each keeps the longest-first order)
    size_t i0 = 0;
    while (i0 < items.size()) {
      long long bytes = 0;
      size_t i1 = i0;
      std::vector<int> seen;
      while (i1 < items.size()) {
        const bool newjob = std::find(seen.begin(), seen.end(), items[i1].job) == seen.end();
        const long long add = items[i1].cnt * wave_b + (newjob ? slot_b : 0);
        if (i1 > i0 && bytes + add > budget) break;
        bytes += add;
        if (newjob) seen.push_back(items[i1].job);
        ++i1;
      }
      std::vector<Item> batch(items.begin() + i0, items.begin() + i1);
      st_.peak = std::max(st_.peak, (double)bytes);
      if (fd) run_fd(batch);
      else run_conv(batch);
      i0 = i1;
    }
  }

  // conventional jobs of several grid sizes (256..2048, band-pruned), one
  // class per size; per slice one potential, one row and one column launch
  // cover every running job / wave of every size
  void run_group(const std::vector<std::pair<std::vector<int>, msl_sim*>>& classes) {
    const long long budget = p_.budget_bytes > 0 ? p_.budget_bytes : (32LL << 30);
    std::vector<GItem> items;
    for (size_t c = 0; c < classes.size(); ++c) {
      msl_sim* cs = classes[c].second;
      use_class(cs, false);
      require(prune_ && kGroupMaxSeg >= (int)classes.size(), MSL_ERR_UNSUPPORTED, "size class cannot be grouped");
      ensure_prop(cs, p_.dz);
      const long long wave_b = np_ * 8;
      const long long maxw = std::max<long long>(1, (budget - wave_b) / (2 * wave_b));
      for (int j : classes[c].first) {
        const msl_job* d = jobs_[j].d;
        if (d->nprobe <= 0) {
          items.push_back({(int)c, {j, 0, 1, d->nslices}});
        } else {
          for (int p0 = 0; p0 < d->nprobe; p0 += (int)maxw)
            items.push_back({(int)c, {j, p0, (int)std::min<long long>(maxw, d->nprobe - p0), d->nslices}});
        }
      }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const GItem& a, const GItem& b) { return a.it.nslices > b.it.nslices; });
    size_t i0 = 0;
    while (i0 < items.size()) {
      long long bytes = 0;
      size_t i1 = i0;
      std::vector<int> seen;
      while (i1 < items.size()) {
        const msl_sim* cs = classes[items[i1].cls].second;
        const long long np = (long long)cs->resX * cs->resX;
        const bool newjob = std::find(seen.begin(), seen.end(), items[i1].it.job) == seen.end();
        const long long add = items[i1].it.cnt * np * 8 + (newjob ? np * 8 : 0);
        if (i1 > i0 && bytes + add > budget) break;
        bytes += add;
        if (newjob) seen.push_back(items[i1].it.job);
        ++i1;
      }
      st_.peak = std::max(st_.peak, (double)bytes);
      run_group_batch(classes, std::vector<GItem>(items.begin() + i0, items.begin() + i1));
      i0 = i1;
    }
  }

 private:
  void run_group_batch(const std::vector<std::pair<std::vector<int>, msl_sim*>>& classes,
                       const std::vector<GItem>& batch) {
    const int nc = (int)classes.size();
    std::vector<Seg> seg(nc);
    for (const GItem& g : batch) seg[g.cls].items.push_back(g.it);
    // job slots are per size (each size's transmissions in its own array); a
    // global job order, longest first, makes the running jobs of every slice
    // a prefix of the potential table
    std::vector<std::pair<int, int>> jorder;  // (class, job), longest first
    for (int c = 0; c < nc; ++c) {
      Seg& S = seg[c];
      S.cs = classes[c].second;
      if (S.items.empty()) continue;
      use_class(S.cs, false);
      bool probes = false;
      for (const Item& it : S.items) probes = probes || jobs_[it.job].d->nprobe > 0;
      if (probes) {
        probe_norm();
        S.norm = norm_;
      }
      scratch_.ensure(np_);
      layout(S.items, S.L);
      for (int j : S.L.jobs) jorder.push_back({c, j});
      S.W.ensure((size_t)S.L.nw * np_);
      S.T.ensure(S.L.jobs.size() * (size_t)np_);
      S.partial.ensure((size_t)S.L.nw * ntiles_);
      S.adf.ensure(S.L.nw);
      for (size_t k = 0; k < S.items.size(); ++k) {
        if (jobs_[S.items[k].job].d->nprobe > 0) continue;
        float2* w = S.W.p + (size_t)S.L.w0[k] * np_;
        HIPC(launch_fill(w, np_, make_float2(1.f, 0.f), st()));
        RowArgs r = row_args(S.cs, w, w, 1, np_);
        r.scale = 1.0f / (float)n_;
        row(S.cs, ROW_FWD, RM_NONE, r);
      }
      norm_ = S.norm;
      init_probes_kxy(S.items, S.L, S.W.p);
    }
    std::stable_sort(jorder.begin(), jorder.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
      return jobs_[a.second].d->nslices > jobs_[b.second].d->nslices;
    });
    // potential table: job k of jorder writes its size's transmission slot
    const int nj = (int)jorder.size();
    std::vector<PotArgs> ptab(nj);
    std::vector<long long> pblk(nj + 1, 0);
    for (int k = 0; k < nj; ++k) {
      Seg& S = seg[jorder[k].first];
      Job& jb = jobs_[jorder[k].second];
      const long long np = (long long)S.cs->resX * S.cs->resX;
      ptab[k] = pot_args(jb.s.get(), jb.s->potmode, 0, 1, p_.dz, 0.f, 0, S.T.p + (size_t)jb.slot * np, np);
      pblk[k + 1] = pblk[k] + potential_tiles(S.cs->resX, S.cs->resX);
    }
    const int ns = batch.front().it.nslices;
    // per-slice pass tables: the running waves of every size
    std::vector<PassGroup<RowArgs>> rg(ns);
    std::vector<PassGroup<ColArgs>> cg(ns);
    for (int i = 1; i <= ns; ++i) {
      RowArgs ra[kGroupMaxSeg];
      ColArgs ca[kGroupMaxSeg];
      int k = 0;
      for (int c = 0; c < nc; ++c) {
        Seg& S = seg[c];
        const long long np = (long long)S.cs->resX * S.cs->resX;
        const int A = S.items.empty() ? 0 : active_waves(S.items, S.L, i);
        ra[k] = row_args(S.cs, S.W.p, S.W.p, A, np);
        ra[k].t = S.T.p;
        ra[k].t_index = S.L.tidx.p;
        ra[k].t_bstride = np;
        ra[k].prune_in = i > 1;
        ra[k].prune_out = 1;
        ca[k] = col_args(S.cs, S.W.p, S.W.p, A, np);
        ca[k].scale = S.cs->nf * S.cs->nf;
        ca[k].tiles_lo = S.cs->prune_tiles;
        ++k;
      }
      HIPC(build_row_group(ra, k, rg[i - 1]));
      HIPC(build_col_group(ca, k, cg[i - 1]));
    }
    DBuf<PotArgs> dptab;
    DBuf<long long> dpblk;
    DBuf<PassGroup<RowArgs>> drg;
    DBuf<PassGroup<ColArgs>> dcg;
    dptab.ensure(nj);
    dpblk.ensure(nj + 1);
    drg.ensure(ns);
    dcg.ensure(ns);
    HIPC(hipMemcpyAsync(dptab.p, ptab.data(), nj * sizeof(PotArgs), hipMemcpyHostToDevice, st()));
    HIPC(hipMemcpyAsync(dpblk.p, pblk.data(), (nj + 1) * sizeof(long long), hipMemcpyHostToDevice, st()));
    HIPC(hipMemcpyAsync(drg.p, rg.data(), ns * sizeof(PassGroup<RowArgs>), hipMemcpyHostToDevice, st()));
    HIPC(hipMemcpyAsync(dcg.p, cg.data(), ns * sizeof(PassGroup<ColArgs>), hipMemcpyHostToDevice, st()));
    HIPC(hipStreamSynchronize(st()));  // host tables may go once copied
    const double t0 = now_ms();
    for (int i = 1; i <= ns; ++i) {
      int ja = 0;
      while (ja < nj && jobs_[jorder[ja].second].d->nslices >= i) ++ja;
      HIPC(launch_potential_group(dptab.p, dpblk.p, ja, pblk[ja], i, st()));
      HIPC(launch_row_group(ROW_INV_MUL_FWD, RM_T, rg[i - 1], drg.p + (i - 1), st()));
      HIPC(launch_col_group(COL_FWD_MUL_INV, CM_PROP, cg[i - 1], dcg.p + (i - 1), st()));
      st_.launches += 3;
      for (int c = 0; c < nc; ++c) {
        Seg& S = seg[c];
        if (S.items.empty()) continue;
        st_.wave_slices += active_waves(S.items, S.L, i);
        use_class(S.cs, false);
        for (size_t k = 0; k < S.items.size(); ++k)
          if (S.items[k].nslices == i) finish_conv(S.items, S.L, k, S.W.p, S.partial.p, S.adf.p);
      }
    }
    HIPC(hipStreamSynchronize(st()));
    st_.prop += now_ms() - t0;
  }

  msl_server* srv_;
  const msl_serve_params& p_;
  std::vector<Job>& jobs_;
  Stats& st_;
  msl_sim* cs_ = nullptr;
  bool fd_ = false, prune_ = false;
  bool fd_prune_ = false;  // finite differences: band limit |k| <= kmax leaves N/4 <= x < 3N/4 empty
  DBuf<float2> Z1_;
  int n_ = 0, ntiles_ = 0;
  long long np_ = 0;
  float norm_ = 1.f;
  DBuf<float2> scratch_;
  DBuf<float> red_;

  hipStream_t st() const { return srv_->stream; }

  // sum |psi| of the probe at the origin (U.cu:2698-2705); integer shifts keep it
  void probe_norm() {
    scratch_.ensure(np_);
    red_.ensure(4096);
    ProbeArgs p = probe_args(scratch_.p, 1);
    HIPC(launch_probe_kspace(p, st()));
    ColArgs c = col_args(cs_, scratch_.p, scratch_.p, 1, np_);
    col(cs_, COL_INV, CM_SCALE, c);
    row(cs_, ROW_INV, RM_NONE, row_args(cs_, scratch_.p, scratch_.p, 1, np_));
    HIPC(launch_l1_normalise(scratch_.p, np_, 1, np_, red_.p, st()));
    norm_ = download_scalar(cs_, red_.p + (np_ + 4095) / 4096);
    require(norm_ > 0 && std::isfinite(norm_), MSL_ERR_RANGE, "probe has zero amplitude (aperture too small?)");
  }

  ProbeArgs probe_args(float2* out, int count) const {
    ProbeArgs p{};
    p.out = out; p.bstride = np_; p.nx = n_; p.ny = n_; p.nprobe = count; p.kx = cs_->kx.p; p.ky = cs_->ky.p;
    p.apert = (float)((double)p_.aperture * 0.001 / (double)cs_->wavel);
    p.ps = cs_->ps; p.df = 0.f; p.Cs = 0.f; p.wavel = cs_->wavel;
    return p;
  }

  // per-batch bookkeeping: wave offsets, job slots, positions, slot index
  struct Layout {
    std::vector<int> w0;        // first wave of each item
    std::vector<int> jobs;      // jobs in slot order (first appearance)
    int nw = 0;
    DBuf<int> tidx, pos;
  };
  // one size inside a grouped batch
  struct Seg {
    msl_sim* cs = nullptr;
    std::vector<Item> items;  // longest first
    Layout L;
    DBuf<float2> W, T;
    DBuf<float> partial, adf;
    float norm = 1.f;
  };

  void layout(const std::vector<Item>& b, Layout& L) {
    L.w0.resize(b.size());
    for (const Item& it : b) jobs_[it.job].slot = -1;
    std::vector<int> tidx, px, py;
    for (size_t k = 0; k < b.size(); ++k) {
      Job& J = jobs_[b[k].job];
      if (J.slot < 0) {
        J.slot = (int)L.jobs.size();
        L.jobs.push_back(b[k].job);
      }
      L.w0[k] = L.nw;
      for (int q = 0; q < b[k].cnt; ++q) {
        tidx.push_back(J.slot);
        px.push_back(J.d->nprobe > 0 ? J.d->posx[b[k].p0 + q] : 0);
        py.push_back(J.d->nprobe > 0 ? J.d->posy[b[k].p0 + q] : 0);
      }
      L.nw += b[k].cnt;
    }
    L.tidx.ensure(L.nw);
    L.pos.ensure(2 * (size_t)L.nw);
    HIPC(hipMemcpyAsync(L.tidx.p, tidx.data(), L.nw * sizeof(int), hipMemcpyHostToDevice, st()));
    HIPC(hipMemcpyAsync(L.pos.p, px.data(), L.nw * sizeof(int), hipMemcpyHostToDevice, st()));
    HIPC(hipMemcpyAsync(L.pos.p + L.nw, py.data(), L.nw * sizeof(int), hipMemcpyHostToDevice, st()));
    HIPC(hipStreamSynchronize(st()));
    st_.batches += 1;
    st_.waves += L.nw;
  }

  // waves still running at slice i (a prefix: items are longest first)
  static int active_waves(const std::vector<Item>& b, const Layout& L, int i) {
    int a = 0;
    for (size_t k = 0; k < b.size(); ++k)
      if (b[k].nslices >= i) a = L.w0[k] + b[k].cnt;
    return a;
  }
  static int active_jobs(const std::vector<Item>& b, const std::vector<Job>& jobs, int i) {
    int a = 0;
    for (const Item& it : b)
      if (it.nslices >= i) a = std::max(a, jobs[it.job].slot + 1);
    return a;
  }

  // probes (STEM items) in the (kx, y) domain: IF_x of the result is the
  // L1-normalised probe of U.cu:2672-2705
  void init_probes_kxy(const std::vector<Item>& b, const Layout& L, float2* W) {
    for (size_t k = 0; k < b.size(); ++k) {
      if (jobs_[b[k].job].d->nprobe <= 0) continue;
      float2* w = W + (size_t)L.w0[k] * np_;
      ProbeArgs p = probe_args(w, b[k].cnt);
      p.posx = L.pos.p + L.w0[k];
      p.posy = L.pos.p + L.nw + L.w0[k];
      HIPC(launch_probe_kspace(p, st()));
      ColArgs c = col_args(cs_, w, w, b[k].cnt, np_);
      c.scale = 1.0f / norm_;
      col(cs_, COL_INV, CM_SCALE, c);
    }
  }

  void adf_out(const std::vector<Item>& b, size_t k, const float* adf_dev) {
    const Job& J = jobs_[b[k].job];
    if (!J.d->adf_out) return;
    HIPC(hipMemcpyAsync(J.d->adf_out + b[k].p0, adf_dev, b[k].cnt * sizeof(float), hipMemcpyDeviceToHost, st()));
  }

  // conventional multislice (U.cu:2085-2124) over one batch
  void run_conv(const std::vector<Item>& b) {
    Layout L;
    layout(b, L);
    const int J = (int)L.jobs.size();
    DBuf<float2> W, T;
    DBuf<float> partial, adf;
    W.ensure((size_t)L.nw * np_);
    T.ensure((size_t)J * np_);
    partial.ensure((size_t)L.nw * ntiles_);
    adf.ensure(L.nw);
    // plane waves: psi = 1, i.e. delta(kx) in the (kx, y) domain
    for (size_t k = 0; k < b.size(); ++k) {
      if (jobs_[b[k].job].d->nprobe > 0) continue;
      float2* w = W.p + (size_t)L.w0[k] * np_;
      HIPC(launch_fill(w, np_, make_float2(1.f, 0.f), st()));
      RowArgs r = row_args(cs_, w, w, 1, np_);
      r.scale = 1.0f / (float)n_;
      row(cs_, ROW_FWD, RM_NONE, r);
    }
    init_probes_kxy(b, L, W.p);
    const double t0 = now_ms();
    for (int i = 1; i <= b.front().nslices; ++i) {
      const int A = active_waves(b, L, i);
      for (int j : L.jobs) {
        Job& jb = jobs_[j];
        if (jb.d->nslices >= i) potential(jb.s.get(), jb.s->potmode, i, 1, p_.dz, 0.f, 0, T.p + (size_t)jb.slot * np_, np_);
      }
      RowArgs r = row_args(cs_, W.p, W.p, A, np_);
      r.t = T.p;
      r.t_index = L.tidx.p;
      r.t_bstride = np_;
      r.prune_in = prune_ && i > 1;
      r.prune_out = prune_;
      row(cs_, ROW_INV_MUL_FWD, RM_T, r);
      ColArgs c = col_args(cs_, W.p, W.p, A, np_);
      c.scale = cs_->nf * cs_->nf;
      c.tiles_lo = prune_ ? cs_->prune_tiles : 0;
      col(cs_, COL_FWD_MUL_INV, CM_PROP, c);
      st_.launches += 2;
      st_.wave_slices += A;
      // items ending here: detector (probes) or exit wave (plane wave)
      for (size_t k = 0; k < b.size(); ++k)
        if (b[k].nslices == i) finish_conv(b, L, k, W.p, partial.p, adf.p);
    }
    HIPC(hipStreamSynchronize(st()));
    st_.prop += now_ms() - t0;
  }

  // item k of a conventional batch after its last slice: detector (probes)
  // or exit wave (plane wave); W in the (kx, y) domain
  void finish_conv(const std::vector<Item>& b, const Layout& L, size_t k, float2* W, float* partial, float* adf) {
    const int det_tiles = prune_ ? 2 * cs_->prune_tiles : ntiles_;
    const Job& jb = jobs_[b[k].job];
    float2* w = W + (size_t)L.w0[k] * np_;
    if (jb.d->nprobe > 0) {
      // unitary |Psi|^2 = |F_y z|^2 for z in the (kx, y) domain
      ColArgs dc = col_args(cs_, w, nullptr, b[k].cnt, np_);
      dc.kin = p_.adf_inner * 1e-3f / cs_->wavel;
      dc.kout = p_.adf_outer * 1e-3f / cs_->wavel;
      dc.det_scale = 1.0f;
      dc.partial = partial + (size_t)L.w0[k] * det_tiles;
      dc.tiles_lo = prune_ ? cs_->prune_tiles : 0;
      col(cs_, COL_FWD_DETECT, CM_SCALE, dc);
      HIPC(launch_reduce_rows(dc.partial, b[k].cnt, det_tiles, adf + L.w0[k], 0, st()));
      adf_out(b, k, adf + L.w0[k]);
    } else if (jb.d->exit_out) {
      RowArgs ri = row_args(cs_, w, scratch_.p, 1, np_);
      ri.prune_in = prune_;
      row(cs_, ROW_INV, RM_NONE, ri);
      HIPC(hipMemcpyAsync(jb.d->exit_out, scratch_.p, np_ * sizeof(float2), hipMemcpyDeviceToHost, st()));
    }
  }

  // finite-difference multislice (U.cu:2040-2083) over one batch
  void run_fd(const std::vector<Item>& b) {
    Layout L;
    layout(b, L);
    const int J = (int)L.jobs.size();
    const size_t wn = (size_t)L.nw * np_;
    DBuf<float2> Pm, P0, Pp, G, Bt, V, Vt;
    DBuf<float> partial, adf;
    Pm.ensure(wn); P0.ensure(wn); Pp.ensure(wn); G.ensure(wn); Bt.ensure(wn);
    if (fd_prune_) Z1_.ensure(wn);
    V.ensure((size_t)J * np_); Vt.ensure((size_t)J * np_);
    partial.ensure((size_t)L.nw * ntiles_);
    adf.ensure(L.nw);
    float2 *pm = Pm.p, *p0 = P0.p, *pp = Pp.p;
    // psi- = probe (real space) or 1; psi = P(dzFD) psi- (InitialiseWavefunctions,
    // STEMInitialiseWavefunctions with the FD start)
    for (size_t k = 0; k < b.size(); ++k)
      if (jobs_[b[k].job].d->nprobe <= 0)
        HIPC(launch_fill(pm + (size_t)L.w0[k] * np_, np_, make_float2(1.f, 0.f), st()));
    init_probes_kxy(b, L, pm);
    for (size_t k = 0; k < b.size(); ++k) {
      if (jobs_[b[k].job].d->nprobe <= 0) continue;
      float2* w = pm + (size_t)L.w0[k] * np_;
      row(cs_, ROW_INV, RM_NONE, row_args(cs_, w, w, b[k].cnt, np_));
    }
    const float nf2 = cs_->nf * cs_->nf;
    auto xform = [&](float2* src, float2* tmp, float2* dst, int nb, int mul) {
      row(cs_, ROW_FWD, RM_NONE, row_args(cs_, src, tmp, nb, np_));
      ColArgs c = col_args(cs_, tmp, tmp, nb, np_);
      c.scale = nf2;
      col(cs_, COL_FWD_MUL_INV, mul, c);
      row(cs_, ROW_INV, RM_NONE, row_args(cs_, tmp, dst, nb, np_));
      st_.launches += 3;
    };
    xform(pm, p0, p0, L.nw, CM_PROP);
    const double t0 = now_ms();
    const float dzfd = cs_->dzFD;
    // Band-limited grids: the step runs fused.  Y holds the Laplacian of psi
    // in the (kx, y) domain.  Per step: one row pass (IF_x of Y, the
    // finite-difference update, F_x of psi+ into Z), two column passes over the
    // in-band columns of Z (band limit -> Z1, band limit times Laplacian -> Y
    // for the next step) and one row pass IF_x(Z1) -> band-limited psi+: four
    // passes instead of the three transforms and the update of U.cu:2040-2083.
    if (fd_prune_) {
      float2 *Y = G.p, *Z = Bt.p, *Z1 = Z1_.p;
      row(cs_, ROW_FWD, RM_NONE, row_args(cs_, p0, Y, L.nw, np_));
      {
        ColArgs c = col_args(cs_, Y, Y, L.nw, np_);
        c.scale = nf2;
        col(cs_, COL_FWD_MUL_INV, CM_GRAD, c);
      }
      for (int it = 1; it <= b.front().nslices; ++it) {
        const int A = active_waves(b, L, it), Ja = active_jobs(b, jobs_, it);
        for (int j : L.jobs) {
          Job& jb = jobs_[j];
          if (jb.d->nslices < it) continue;
          const float zc = (jb.s->MaxZ - jb.s->MinZ) - dzfd * (float)it;
          potential(jb.s.get(), 2, 0, 1, 1.0f, zc, 1, V.p + (size_t)jb.slot * np_, np_);
        }
        xform(V.p, Vt.p, V.p, Ja, CM_BAND);  // band-limited V
        RowArgs r = row_args(cs_, Y, Z, A, np_);
        r.prune_in = r.prune_out = 1;
        r.t = V.p;
        r.t_index = L.tidx.p;
        r.t_bstride = np_;
        r.fd_pm = pm;
        r.fd_p0 = p0;
        r.fd_dz = dzfd;
        r.fd_wavel = cs_->wavel;
        r.sigma = cs_->sigma2;
        row(cs_, ROW_INV_FD_FWD, RM_T, r);
        ColArgs c = col_args(cs_, Z, Z1, A, np_);
        c.scale = nf2;
        c.tiles_lo = cs_->prune_tiles;
        col(cs_, COL_FWD_MUL_INV, CM_BAND, c);
        c.dst = Y;
        col(cs_, COL_FWD_MUL_INV, CM_BANDGRAD, c);
        RowArgs ri = row_args(cs_, Z1, pp, A, np_);
        ri.prune_in = 1;
        row(cs_, ROW_INV, RM_NONE, ri);
        st_.launches += 4;
        std::swap(pm, p0);
        std::swap(p0, pp);
        st_.wave_slices += A;
        for (size_t k = 0; k < b.size(); ++k)
          if (b[k].nslices == it) finish_fd(b, L, k, p0, Z, partial.p, adf.p);
      }
      HIPC(hipStreamSynchronize(st()));
      st_.prop += now_ms() - t0;
      return;
    }
    for (int it = 1; it <= b.front().nslices; ++it) {
      const int A = active_waves(b, L, it), Ja = active_jobs(b, jobs_, it);
      for (int j : L.jobs) {
        Job& jb = jobs_[j];
        if (jb.d->nslices < it) continue;
        const float zc = (jb.s->MaxZ - jb.s->MinZ) - dzfd * (float)it;
        potential(jb.s.get(), 2, 0, 1, 1.0f, zc, 1, V.p + (size_t)jb.slot * np_, np_);
      }
      xform(V.p, Vt.p, V.p, Ja, CM_BAND);  // band-limited V
      xform(p0, G.p, G.p, A, CM_GRAD);     // spectral Laplacian
      for (size_t k = 0; k < b.size(); ++k) {
        if (b[k].nslices < it) continue;
        const size_t o = (size_t)L.w0[k] * np_;
        HIPC(launch_fd_update(G.p + o, pm + o, pp + o, p0 + o, V.p + (size_t)jobs_[b[k].job].slot * np_,
                              (long long)b[k].cnt * np_, np_, cs_->sigma2, cs_->wavel, dzfd, st()));
      }
      xform(pp, Bt.p, pp, A, CM_BAND);     // band-limited psi+
      std::swap(pm, p0);
      std::swap(p0, pp);
      st_.wave_slices += A;
      for (size_t k = 0; k < b.size(); ++k)
        if (b[k].nslices == it) finish_fd(b, L, k, p0, Bt.p, partial.p, adf.p);
    }
    HIPC(hipStreamSynchronize(st()));
    st_.prop += now_ms() - t0;
  }

  // item k of a finite-difference batch after its last step: detector
  // (probes, through scratch waves f) or exit wave; W in real space
  void finish_fd(const std::vector<Item>& b, const Layout& L, size_t k, float2* W, float2* f, float* partial,
                 float* adf) {
    const Job& jb = jobs_[b[k].job];
    float2* w = W + (size_t)L.w0[k] * np_;
    if (jb.d->nprobe > 0) {
      float2* fw = f + (size_t)L.w0[k] * np_;
      row(cs_, ROW_FWD, RM_NONE, row_args(cs_, w, fw, b[k].cnt, np_));
      ColArgs dc = col_args(cs_, fw, nullptr, b[k].cnt, np_);
      dc.kin = p_.adf_inner * 1e-3f / cs_->wavel;
      dc.kout = p_.adf_outer * 1e-3f / cs_->wavel;
      dc.det_scale = cs_->nf * cs_->nf;
      dc.partial = partial + (size_t)L.w0[k] * ntiles_;
      col(cs_, COL_FWD_DETECT, CM_SCALE, dc);
      HIPC(launch_reduce_rows(dc.partial, b[k].cnt, ntiles_, adf + L.w0[k], 0, st()));
      adf_out(b, k, adf + L.w0[k]);
    } else if (jb.d->exit_out) {
      HIPC(hipMemcpyAsync(jb.d->exit_out, w, np_ * sizeof(float2), hipMemcpyDeviceToHost, st()));
    }
  }
};

void check_job(const msl_job& d, int i) {
  const std::string w = "job " + std::to_string(i) + ": ";
  require(fft_size_supported(d.n), MSL_ERR_UNSUPPORTED, w + "grid must be a power of two in 16..2048");
  require(d.method == MSL_METHOD_CONVENTIONAL || d.method == MSL_METHOD_FD, MSL_ERR_ARG, w + "unknown method");
  require(d.nslices >= 1, MSL_ERR_ARG, w + "nslices must be positive");
  require(d.natoms >= 1 && d.Z && d.x && d.y && d.z, MSL_ERR_ARG, w + "no atoms");
  require(d.nprobe >= 0, MSL_ERR_ARG, w + "negative probe count");
  // one detector launch per job and batch holds its probes on grid.y
  require(d.nprobe <= 65535, MSL_ERR_UNSUPPORTED, w + "more than 65535 probes in one job");
  if (d.nprobe > 0) {
    require(d.posx && d.posy && d.adf_out, MSL_ERR_ARG, w + "probe positions and adf_out required");
    for (int p = 0; p < d.nprobe; ++p)
      require(d.posx[p] >= 0 && d.posx[p] < d.n && d.posy[p] >= 0 && d.posy[p] < d.n, MSL_ERR_ARG,
              w + "probe position outside the grid");
  }
}

}  // namespace

extern "C" {

msl_status msl_server_create(msl_server** out, int device) {
  if (!out) return MSL_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MSL_ERR_HIP;
  if (device < 0 || device >= ndev) return MSL_ERR_ARG;
  msl_server* s = new (std::nothrow) msl_server;
  if (!s) return MSL_ERR_OOM;
  s->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return MSL_ERR_HIP;
  }
  *out = s;
  return MSL_OK;
}

void msl_server_destroy(msl_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) {
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamDestroy(s->stream);
  }
  delete s;
}

const char* msl_server_last_error(const msl_server* s) { return s ? s->err.c_str() : "null handle"; }

msl_status msl_server_run(msl_server* srv, const msl_serve_params* prm, const msl_job* jd, int njobs,
                          double* stats) {
  if (!srv) return MSL_ERR_ARG;
  try {
    HIPC(hipSetDevice(srv->device));
    require(prm && jd && njobs >= 1, MSL_ERR_ARG, "no jobs");
    require(prm->kV > 0 && prm->ps > 0 && prm->dz > 0, MSL_ERR_ARG, "kV, pixel size and slice thickness must be positive");
    require(prm->potential_mode == MSL_POT_INTEGRATED || prm->potential_mode == MSL_POT_PROJECTED, MSL_ERR_ARG,
            "unknown potential mode");
    for (int i = 0; i < njobs; ++i) check_job(jd[i], i);
    const double t0 = now_ms();
    Stats st;
    // one handle per job on the server's stream: extents, binning, physics
    std::vector<Job> jobs(njobs);
    std::map<std::tuple<int, int, int>, std::vector<int>> classes;
    for (int i = 0; i < njobs; ++i) {
      const msl_job& d = jd[i];
      Job& J = jobs[i];
      J.d = &d;
      J.stem = d.nprobe > 0;
      J.s.reset(new msl_sim);
      msl_sim* s = J.s.get();
      s->device = srv->device;
      s->stream = srv->stream;
      s->own_stream = false;
      s->potmode = prm->potential_mode;
      set_atoms(s, std::vector<int>(d.Z, d.Z + d.natoms), std::vector<float>(d.x, d.x + d.natoms),
                std::vector<float>(d.y, d.y + d.natoms), std::vector<float>(d.z, d.z + d.natoms));
      bin_atoms(s, nullptr, nullptr, nullptr);
      s->kV = prm->kV;
      s->obj = J.stem ? prm->aperture : 0.f;
      s->have_micro = true;
      calc_vars(s, prm->ps, 0.f, d.n, d.n, s->MaxZ - s->MinZ, 0.f, 0.f, J.stem);
      precalc(s);
      ensure_ztab(s);
      if (d.method == MSL_METHOD_FD)
        require(std::isfinite(s->dzFD) && s->dzFD > 0, MSL_ERR_RANGE,
                "job " + std::to_string(i) + ": no stable finite-difference step");
      classes[{d.n, d.method, d.method == MSL_METHOD_FD && J.stem ? 1 : 0}].push_back(i);
    }
    st.setup = now_ms() - t0;
    Runner R(srv, *prm, jobs, st);
    std::vector<std::pair<std::vector<int>, msl_sim*>> grouped;
    for (auto& kv : classes) {
      const int first = kv.second.front();
      msl_sim* cs = jobs[first].s.get();
      const int n = std::get<0>(kv.first);
      const bool fd = std::get<1>(kv.first) == MSL_METHOD_FD;
      if (prm->group_sizes && !fd && cs->prune_tiles > 0 && (n == 256 || n == 512 || n == 1024 || n == 2048))
        grouped.push_back({kv.second, cs});
      else
        R.run_class(kv.second, cs, fd);
    }
    if (!grouped.empty()) R.run_group(grouped);
    HIPC(hipStreamSynchronize(srv->stream));
    if (stats) {
      stats[0] = now_ms() - t0; stats[1] = st.setup; stats[2] = st.prop; stats[3] = st.batches;
      stats[4] = st.waves; stats[5] = st.wave_slices; stats[6] = st.launches; stats[7] = st.peak;
    }
  } catch (const Error& e) {
    (void)hipStreamSynchronize(srv->stream);
    srv->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    (void)hipStreamSynchronize(srv->stream);
    srv->err = "host allocation failed";
    return MSL_ERR_OOM;
  }
  srv->err.clear();
  return MSL_OK;
}

}  // extern "C"
