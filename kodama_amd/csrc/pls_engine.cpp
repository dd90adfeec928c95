// Host side of the device PLS-LDA cross-validation: fold bookkeeping, the
// class-grouped training rows of each evaluation, launches and result checks.
// The fold partition and the evaluation order follow run_plscv_host
// (plscv.cpp:2293-2507): folds in ascending id order, validation rows in
// ascending order, the selected component count = min over folds.
#include "pls_engine.h"

#include <algorithm>
#include <map>
#include <random>
#include <stdexcept>

#include "pls_launch.h"

namespace kodama_hip {

std::vector<double> pls_start_vector(int p) {
  std::mt19937 gen(1u);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::vector<double> w(static_cast<size_t>(p));
  for (double& v : w) v = static_cast<double>(nd(gen));
  return w;
}

struct PlsCv::Impl {
  int device = 0, n = 0, p = 0, classes = 0;
  int requested = 0, max_components = 0, rank_cap = 0;
  bool center = true, scale = true, da = false;
  Stream stream;
  DeviceBuffer<float> x;
  DeviceBuffer<double> w0, mean, sd, xsum, gram, weights, loadings, gw, zsum, mu, linear, constant, val_scores, disc;
  DeviceBuffer<double> fsum;
  DeviceBuffer<int> seg_code, seg_count, val_rows, val_off, run_first, run_end, fcnt;
  DeviceBuffer<int> shift_row, chunk_fold, chunk_begin, chunk_end, chunk_off;
  DeviceBuffer<double> moment_partial, moments;
  DeviceBuffer<int> codes, pred, comps, rank;
  IVec groups;  // row ranges of the runs: [groups[g], groups[g+1])
  IVec cached_folds, fold_ids, h_val_rows, h_val_off;
  IVec h_comps, h_rank;
  PinnedBuffer<int> h_codes, h_pred, h_small;  // codes, predictions; components and ranks
  EventTimer timer;
  double eval_ms_total = 0.0;
  int eval_calls = 0;

  explicit Impl(int dev) : device(dev), stream(dev) {}

  void rebuild_folds(const IVec& folds);
};

void PlsCv::Impl::rebuild_folds(const IVec& folds) {
  // A fold belongs to the group (run) of its rows; its training rows are the
  // other rows of that group only.
  std::map<int, int> group_of;
  for (int g = 0; g + 1 < static_cast<int>(groups.size()); ++g) {
    for (int i = groups[g]; i < groups[g + 1]; ++i) {
      const auto it = group_of.emplace(folds[i], g).first;
      if (it->second != g) throw InvalidArgument("a fold id spans two runs.");
    }
  }
  fold_ids.clear();
  for (const auto& kv : group_of) fold_ids.push_back(kv.first);
  const int F = static_cast<int>(fold_ids.size());
  // Each fold's training sums are its run's other folds' sums: the folds of a
  // run must be consecutive in fold-id order.
  IVec h_first(static_cast<size_t>(F)), h_end(static_cast<size_t>(F));
  for (int f = 0; f < F;) {
    const int g = group_of[fold_ids[f]];
    int e = f + 1;
    while (e < F && group_of[fold_ids[e]] == g) ++e;
    for (int t = f; t < e; ++t) {
      h_first[t] = f;
      h_end[t] = e;
    }
    f = e;
  }
  {
    std::vector<char> seen(groups.size(), 0);
    for (int f = 0; f < F; ++f) {
      const int g = group_of[fold_ids[f]];
      if (h_first[f] == f && seen[g]) throw InvalidArgument("the fold ids of a run must be consecutive.");
      seen[g] = 1;
    }
  }
  h_val_rows.clear();
  h_val_off.assign(1, 0);
  IVec h_shift(static_cast<size_t>(F));
  for (int f = 0; f < F; ++f) {
    const int fid = fold_ids[f];
    const int g = group_of[fid];
    for (int i = groups[g]; i < groups[g + 1]; ++i) {
      if (folds[i] == fid) h_val_rows.push_back(i);
    }
    h_val_off.push_back(static_cast<int>(h_val_rows.size()));
    if (h_val_off[f + 1] - h_val_off[f] == groups[g + 1] - groups[g]) {
      throw std::runtime_error("PLS fold has no training samples.");
    }
    h_shift[f] = groups[g];
  }
  // Row chunks of each fold's own rows for the moment kernel.
  IVec h_cf, h_cb, h_ce, h_co(1, 0);
  for (int f = 0; f < F; ++f) {
    for (int b = h_val_off[f]; b < h_val_off[f + 1]; b += kPlsMomentRows) {
      h_cf.push_back(f);
      h_cb.push_back(b);
      h_ce.push_back(std::min(b + kPlsMomentRows, h_val_off[f + 1]));
    }
    h_co.push_back(static_cast<int>(h_cf.size()));
  }
  const hipStream_t s = stream.get();
  val_rows.resize(h_val_rows.size());
  val_off.resize(h_val_off.size());
  run_first.resize(h_first.size());
  run_end.resize(h_end.size());
  shift_row.resize(h_shift.size());
  chunk_fold.resize(h_cf.size());
  chunk_begin.resize(h_cb.size());
  chunk_end.resize(h_ce.size());
  chunk_off.resize(h_co.size());
  val_rows.upload(h_val_rows.data(), h_val_rows.size(), s);
  val_off.upload(h_val_off.data(), h_val_off.size(), s);
  run_first.upload(h_first.data(), h_first.size(), s);
  run_end.upload(h_end.data(), h_end.size(), s);
  shift_row.upload(h_shift.data(), h_shift.size(), s);
  chunk_fold.upload(h_cf.data(), h_cf.size(), s);
  chunk_begin.upload(h_cb.data(), h_cb.size(), s);
  chunk_end.upload(h_ce.data(), h_ce.size(), s);
  chunk_off.upload(h_co.data(), h_co.size(), s);
  const size_t z16 = static_cast<size_t>((p + 1 + 15) / 16 * 16);
  moment_partial.resize(h_cf.size() * z16 * z16);
  moments.resize(static_cast<size_t>(F) * (p + 1) * (p + 1));
  const size_t fp = static_cast<size_t>(F) * p;
  const size_t fc = static_cast<size_t>(F) * classes;
  mean.resize(fp);
  sd.resize(fp);
  xsum.resize(fp);
  gram.resize(fp * p);
  weights.resize(fp * kPlsMaxComponents);
  fsum.resize(fc * p);
  fcnt.resize(fc);
  comps.resize(static_cast<size_t>(F));
  rank.resize(static_cast<size_t>(F));
  seg_code.resize(fc);
  seg_count.resize(static_cast<size_t>(F));
  if (!pls_eval_compact(p, classes, rank_cap)) {
    // The wide layout's per-fold scratch (the compact one keeps it in LDS).
    loadings.resize(fp * kPlsMaxComponents);
    gw.resize(fp * kPlsMaxComponents);
    zsum.resize(fc * p);
    mu.resize(fc * kPlsMaxComponents);
    linear.resize(fc * kPlsMaxComponents);
    constant.resize(fc);
  }
  PlsCacheParams c;
  c.x = x.get();
  c.p = p;
  c.folds = F;
  c.center = center ? 1 : 0;
  c.scale = scale ? 1 : 0;
  c.val.rows = val_rows.get();
  c.val.offsets = val_off.get();
  c.shift_row = shift_row.get();
  c.run_first = run_first.get();
  c.run_end = run_end.get();
  c.chunks = static_cast<int>(h_cf.size());
  c.chunk_fold = chunk_fold.get();
  c.chunk_begin = chunk_begin.get();
  c.chunk_end = chunk_end.get();
  c.chunk_off = chunk_off.get();
  c.partial = moment_partial.get();
  c.moments = moments.get();
  c.mean = mean.get();
  c.sd = sd.get();
  c.xsum = xsum.get();
  c.gram = gram.get();
  check(launch_pls_cache(c, s), "pls_cache launch");
  cached_folds = folds;
}

PlsCv::PlsCv(int device, const float* x, int n, int p, const kodama_hip_core_options& o, int classes,
             const IVec& run_offsets) {
  if (o.max_components < 1) throw InvalidArgument("PLSOptions::max_components must be positive.");
  if (p > kPlsMaxDims) throw InvalidArgument("the HIP PLS-LDA kernels support up to 256 feature columns.");
  Impl* m = new Impl(device);
  m->n = n;
  m->p = p;
  m->groups = run_offsets.empty() ? IVec{0, n} : run_offsets;
  if (m->groups.front() != 0 || m->groups.back() != n) {
    delete m;
    throw InvalidArgument("run offsets must cover the rows.");
  }
  m->classes = std::max(1, classes);
  m->max_components = o.max_components;
  m->requested = o.fixed_components > 0 ? o.fixed_components : o.max_components;
  m->rank_cap = std::min(o.max_components, p);
  m->center = o.center != 0;
  m->scale = o.scale != 0;
  if (m->rank_cap > kPlsMaxComponents) {
    delete m;
    throw InvalidArgument("the HIP PLS-LDA kernels support up to 64 components.");
  }
  const hipStream_t s = m->stream.get();
  m->x.resize(static_cast<size_t>(n) * p);
  m->x.upload(x, m->x.size(), s);
  const std::vector<double> w0 = pls_start_vector(p);
  m->w0.resize(w0.size());
  m->w0.upload(w0.data(), w0.size(), s);
  if (static_cast<size_t>(m->classes) * p > static_cast<size_t>(kPlsMaxSLds) ||
      pls_eval_lds_bytes(p, m->classes, m->rank_cap, pls_eval_compact(p, m->classes, m->rank_cap)) > 160 * 1024) {
    delete m;
    throw InvalidArgument("the HIP PLS-LDA kernel holds classes x columns <= 8192 in LDS.");
  }
  m->codes.resize(static_cast<size_t>(n));
  m->pred.resize(static_cast<size_t>(n));
  m->h_codes.resize(static_cast<size_t>(n));
  m->h_pred.resize(static_cast<size_t>(n));
  m->val_scores.resize(static_cast<size_t>(n) * kPlsMaxComponents);
  m->disc.resize(static_cast<size_t>(n) * m->classes);
  impl_ = m;
}

PlsCv::~PlsCv() { delete impl_; }

void PlsCv::evaluate(const IVec& codes, const IVec& folds, int* pred_codes, int* selected) {
  Impl& m = *impl_;
  if (static_cast<int>(codes.size()) != m.n) throw InvalidArgument("labels and folds must have one entry per row.");
  std::copy(codes.begin(), codes.end(), m.h_codes.get());
  evaluate_staged(folds, selected);
  std::copy(m.h_pred.get(), m.h_pred.get() + m.n, pred_codes);
}

int* PlsCv::staged_codes() { return impl_->h_codes.get(); }
const int* PlsCv::staged_pred() const { return impl_->h_pred.get(); }

void PlsCv::prepare(const IVec& folds) {
  Impl& m = *impl_;
  if (static_cast<int>(folds.size()) != m.n) throw InvalidArgument("labels and folds must have one entry per row.");
  if (folds != m.cached_folds) m.rebuild_folds(folds);
  m.stream.sync();
}

void PlsCv::evaluate_staged(const IVec& folds, int* selected) {
  Impl& m = *impl_;
  if (static_cast<int>(folds.size()) != m.n) throw InvalidArgument("labels and folds must have one entry per row.");
  const int* codes = m.h_codes.get();
  bool in_range = true;
  for (int i = 0; i < m.n; ++i) in_range &= codes[i] >= 0 && codes[i] < m.classes;
  if (!in_range) throw InvalidArgument("label code out of range.");
  if (folds != m.cached_folds) m.rebuild_folds(folds);
  const int F = static_cast<int>(m.fold_ids.size());
  const hipStream_t s = m.stream.get();
  m.codes.upload(codes, static_cast<size_t>(m.n), s);
  PlsFoldSumsParams fsp;
  fsp.x = m.x.get();
  fsp.p = m.p;
  fsp.folds = F;
  fsp.max_classes = m.classes;
  fsp.codes = m.codes.get();
  fsp.val.rows = m.val_rows.get();
  fsp.val.offsets = m.val_off.get();
  fsp.fsum = m.fsum.get();
  fsp.fcnt = m.fcnt.get();
  PlsEvalParams e;
  e.x = m.x.get();
  e.p = m.p;
  e.folds = F;
  e.requested = m.requested;
  e.da = m.da ? 1 : 0;
  e.rank_cap = m.rank_cap;
  e.max_classes = m.classes;
  e.w0 = m.w0.get();
  e.mean = m.mean.get();
  e.sd = m.sd.get();
  e.xsum = m.xsum.get();
  e.gram = m.gram.get();
  e.run_first = m.run_first.get();
  e.run_end = m.run_end.get();
  e.fsum = m.fsum.get();
  e.fcnt = m.fcnt.get();
  e.val.rows = m.val_rows.get();
  e.val.offsets = m.val_off.get();
  e.seg_code = m.seg_code.get();
  e.seg_count = m.seg_count.get();
  e.weights = m.weights.get();
  e.loadings = m.loadings.get();
  e.gw = m.gw.get();
  e.zsum = m.zsum.get();
  e.mu = m.mu.get();
  e.linear = m.linear.get();
  e.constant = m.constant.get();
  e.pred_code = m.pred.get();
  e.components = m.comps.get();
  e.rank = m.rank.get();
  e.val_scores = m.val_scores.get();
  e.discriminants = m.disc.get();
  m.timer.start(s);
  check(launch_pls_fold_sums(fsp, s), "pls_fold_sums launch");
  check(launch_pls_eval(e, s), "pls_eval launch");
  m.timer.stop(s);
  if (m.h_small.size() < 2 * static_cast<size_t>(F)) m.h_small.resize(2 * static_cast<size_t>(F));
  m.pred.download(m.h_pred.get(), static_cast<size_t>(m.n), s);
  m.comps.download(m.h_small.get(), static_cast<size_t>(F), s);
  m.rank.download(m.h_small.get() + F, static_cast<size_t>(F), s);
  m.stream.sync();
  m.h_comps.assign(m.h_small.get(), m.h_small.get() + F);
  m.h_rank.assign(m.h_small.get() + F, m.h_small.get() + 2 * F);
  m.eval_ms_total += m.timer.ms();
  ++m.eval_calls;
  int evaluated = m.requested;
  for (int f = 0; f < F; ++f) {
    if (m.h_comps[f] < 0) throw std::runtime_error("PLS-LDA Cholesky factorization failed.");
    evaluated = std::min(evaluated, m.h_comps[f]);
  }
  if (selected) *selected = std::min(evaluated, m.max_components);
}

void PlsCv::set_da(bool da) { impl_->da = da; }

double PlsCv::eval_kernel_ms_average() const {
  return impl_->eval_calls ? impl_->eval_ms_total / impl_->eval_calls : 0.0;
}

PlsFoldDetail PlsCv::detail(int fold) const {
  const Impl& m = *impl_;
  const int F = static_cast<int>(m.fold_ids.size());
  if (fold < 0 || fold >= F || m.h_comps.empty()) throw InvalidArgument("no such fold in the last evaluation.");
  PlsFoldDetail d;
  d.rank = m.h_rank[fold];
  d.components = m.h_comps[fold];
  {
    int count = 0;
    m.seg_count.download(&count, 1, m.stream.get(), static_cast<size_t>(fold));
    d.classes.assign(static_cast<size_t>(std::max(0, count)), 0);
    m.seg_code.download(d.classes.data(), d.classes.size(), m.stream.get(), static_cast<size_t>(fold) * m.classes);
    check(hipStreamSynchronize(m.stream.get()), "detail sync");
  }
  d.val_rows.assign(m.h_val_rows.begin() + m.h_val_off[fold], m.h_val_rows.begin() + m.h_val_off[fold + 1]);
  const int nv = static_cast<int>(d.val_rows.size());
  const int cf = static_cast<int>(d.classes.size());
  const hipStream_t s = m.stream.get();
  std::vector<double> w(static_cast<size_t>(m.p) * kPlsMaxComponents);  // [a * p + j]
  std::vector<double> tv(static_cast<size_t>(nv) * kPlsMaxComponents);
  std::vector<double> disc(static_cast<size_t>(nv) * m.classes);
  m.weights.download(w.data(), w.size(), s, static_cast<size_t>(fold) * m.p * kPlsMaxComponents);
  m.val_scores.download(tv.data(), tv.size(), s, static_cast<size_t>(m.h_val_off[fold]) * kPlsMaxComponents);
  m.disc.download(disc.data(), disc.size(), s, static_cast<size_t>(m.h_val_off[fold]) * m.classes);
  check(hipStreamSynchronize(s), "detail sync");
  const bool fitted = d.rank > 0 && cf > 1;
  const int A = fitted ? d.components : 0;
  d.weights.assign(static_cast<size_t>(m.p) * (fitted ? d.rank : 0), 0.0);
  for (int j = 0; j < m.p && fitted; ++j)
    for (int a = 0; a < d.rank; ++a) d.weights[static_cast<size_t>(j) * d.rank + a] = w[static_cast<size_t>(a) * m.p + j];
  d.val_scores.assign(static_cast<size_t>(nv) * A, 0.0);
  d.discriminants.assign(static_cast<size_t>(nv) * (fitted ? cf : 0), 0.0);
  for (int i = 0; i < nv && fitted; ++i) {
    for (int a = 0; a < A; ++a) d.val_scores[static_cast<size_t>(i) * A + a] = tv[static_cast<size_t>(i) * kPlsMaxComponents + a];
    for (int c = 0; c < cf; ++c) d.discriminants[static_cast<size_t>(i) * cf + c] = disc[static_cast<size_t>(i) * m.classes + c];
  }
  return d;
}

PlsGeneralResult pls_general_fit(int device, const float* x, const float* y, int n, int p, int q,
                                 const kodama_hip_core_options& o) {
  const int requested = o.fixed_components > 0 ? o.fixed_components : o.max_components;
  if (requested < 1) throw InvalidArgument("PLS component count must be positive.");
  if (p > kPlsMaxDims) throw InvalidArgument("the HIP PLS kernels support up to 256 predictors.");
  const int rank_cap = std::max(1, std::min({requested, p, std::max(1, n - 1)}));
  if (rank_cap > kPlsMaxComponents) throw InvalidArgument("the HIP PLS kernels support up to 64 components.");
  if (static_cast<size_t>(q) * p > static_cast<size_t>(kPlsMaxSLds) ||
      pls_fit_lds_bytes(p, q, rank_cap, false) > 160 * 1024 ||
      (p + q + 15) / 16 * 16 * 18 * sizeof(double) > 64 * 1024) {
    throw InvalidArgument("the HIP PLS kernels hold responses x predictors <= 8192 and predictors + responses <= 448.");
  }
  Stream st(device);
  const hipStream_t s = st.get();
  DeviceBuffer<float> dx(static_cast<size_t>(n) * p), dy(static_cast<size_t>(n) * q);
  dx.upload(x, dx.size(), s);
  dy.upload(y, dy.size(), s);
  // ym centres Y; yreport is the reference's y_center: the column means when
  // centring, else the column sums (fit_general_pls accumulates them and
  // divides only when centring), which the fitted values start from.
  DeviceBuffer<double> xm(p), xsd(p), ym(q), ysd(q), ysum(q);
  EventTimer t_all, t_cross;
  t_all.start(s);
  // Column statistics in row chunks, reduced in chunk order.
  const int st_chunks = std::max(1, std::min(2048, (n + 63) / 64));
  const int st_rows = (n + st_chunks - 1) / st_chunks;
  DeviceBuffer<double> partial(static_cast<size_t>(st_chunks) * std::max(p, q));
  auto stats = [&](const float* m, int cols, double* mean, double* sd, int scale, double* sum) {
    PlsColStatsParams c;
    c.m = m;
    c.n = n;
    c.cols = cols;
    c.chunks = st_chunks;
    c.chunk_rows = st_rows;
    c.center = o.center != 0 ? 1 : 0;
    c.mean = mean;
    c.sd = sd;
    c.sum = sum;
    c.partial = partial.get();
    c.pass = 0;
    check(launch_pls_colstats(c, s), "pls_colstats launch");
    check(launch_pls_colstats_reduce(c, scale, s), "pls_colstats reduce");
    c.pass = 1;
    c.sum = nullptr;
    if (scale) check(launch_pls_colstats(c, s), "pls_colstats launch");
    check(launch_pls_colstats_reduce(c, scale, s), "pls_colstats reduce");
  };
  stats(dx.get(), p, xm.get(), xsd.get(), o.scale != 0 ? 1 : 0, nullptr);
  stats(dy.get(), q, ym.get(), ysd.get(), 0, ysum.get());
  const double* yreport = o.center != 0 ? ym.get() : ysum.get();
  // Cross-products on the matrix cores.
  const int p16 = (p + 15) / 16 * 16, z16 = (p + q + 15) / 16 * 16;
  const int cr_rows = std::max(16, ((n + 1023) / 1024 + 15) / 16 * 16);
  const int cr_chunks = (n + cr_rows - 1) / cr_rows;
  DeviceBuffer<double> cpart(static_cast<size_t>(cr_chunks) * p16 * z16), gram(static_cast<size_t>(p) * p),
    cross(static_cast<size_t>(q) * p);
  PlsCrossParams cp;
  cp.x = dx.get();
  cp.y = dy.get();
  cp.n = n;
  cp.p = p;
  cp.q = q;
  cp.chunks = cr_chunks;
  cp.chunk_rows = cr_rows;
  cp.x_mean = xm.get();
  cp.x_sd = xsd.get();
  cp.y_mean = ym.get();
  cp.partial = cpart.get();
  cp.gram = gram.get();
  cp.cross = cross.get();
  t_cross.start(s);
  check(launch_pls_cross(cp, s), "pls_cross launch");
  t_cross.stop(s);
  // SIMPLS, response weights, coefficients.
  const std::vector<double> w0 = pls_start_vector(p);
  DeviceBuffer<double> dw0(p), W(static_cast<size_t>(p) * kPlsMaxComponents), V(W.size()),
    Q(static_cast<size_t>(q) * kPlsMaxComponents), B(static_cast<size_t>(p) * q);
  DeviceBuffer<int> dfit(1);
  dw0.upload(w0.data(), w0.size(), s);
  PlsFitParams fp;
  fp.p = p;
  fp.q = q;
  fp.rank_cap = rank_cap;
  fp.w0 = dw0.get();
  fp.gram = gram.get();
  fp.cross = cross.get();
  fp.weights = W.get();
  fp.loadings = V.get();
  fp.yweights = Q.get();
  fp.coefficients = B.get();
  fp.fitted = dfit.get();
  DeviceBuffer<double> wscratch(2 * static_cast<size_t>(kPlsMaxComponents) * p);
  fp.wscratch = wscratch.get();
  check(launch_pls_fit(fp, s), "pls_fit launch");
  int A = 0;
  dfit.download(&A, 1, s);
  st.sync();
  if (A < 1) throw DeviceError("PLS: SIMPLS found no supervised latent direction.");
  DeviceBuffer<float> dsc(static_cast<size_t>(n) * A), dft(static_cast<size_t>(n) * q);
  DeviceBuffer<double> tbuf(pls_scores_staged(p, q, A) ? 0 : static_cast<size_t>(n) * A);
  PlsScoreParams sp;
  sp.x = dx.get();
  sp.n = n;
  sp.p = p;
  sp.q = q;
  sp.A = A;
  sp.x_mean = xm.get();
  sp.x_sd = xsd.get();
  sp.y_mean = yreport;
  sp.weights = W.get();
  sp.yweights = Q.get();
  sp.tbuf = tbuf.get();
  sp.scores = dsc.get();
  sp.fitted = dft.get();
  check(launch_pls_scores(sp, s), "pls_scores launch");
  t_all.stop(s);
  PlsGeneralResult r;
  r.components = A;
  std::vector<double> hw(W.size()), hq(Q.size()), hb(B.size()), hxm(p), hxs(p), hym(q);
  W.download(hw.data(), hw.size(), s);
  Q.download(hq.data(), hq.size(), s);
  B.download(hb.data(), hb.size(), s);
  xm.download(hxm.data(), hxm.size(), s);
  xsd.download(hxs.data(), hxs.size(), s);
  check(hipMemcpyAsync(hym.data(), yreport, sizeof(double) * q, hipMemcpyDeviceToHost, s), "y_center download");
  r.scores.resize(dsc.size());
  r.fitted.resize(dft.size());
  dsc.download(r.scores.data(), r.scores.size(), s);
  dft.download(r.fitted.data(), r.fitted.size(), s);
  st.sync();
  r.weights.resize(static_cast<size_t>(p) * A);
  r.response_loadings.resize(static_cast<size_t>(q) * A);
  for (int j = 0; j < p; ++j)
    for (int a = 0; a < A; ++a) r.weights[static_cast<size_t>(j) * A + a] = static_cast<float>(hw[static_cast<size_t>(j) * kPlsMaxComponents + a]);
  for (int c = 0; c < q; ++c)
    for (int a = 0; a < A; ++a)
      r.response_loadings[static_cast<size_t>(c) * A + a] = static_cast<float>(hq[static_cast<size_t>(c) * kPlsMaxComponents + a]);
  r.coefficients.assign(hb.begin(), hb.end());
  r.x_center.assign(hxm.begin(), hxm.end());
  r.x_scale.assign(hxs.begin(), hxs.end());
  r.y_center.assign(hym.begin(), hym.end());
  r.cross_ms = t_cross.ms();
  r.total_ms = t_all.ms();
  return r;
}

}  // namespace kodama_hip
