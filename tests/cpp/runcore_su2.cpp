// The reference's SU(2) cutoff loop driven from C++ through the C ABI of
// include/spfrg.h, in the shape the reference-side binding takes:
//
//   * HipSU2EffectiveAction -- the state and the flow of one GPU context.  The
//     flow object's isDiverged() asks the device (sp_is_diverged); it is what
//     FrgCore::_flow points to, so SpinParser::runCore's divergence check
//     (src/SpinParser.cpp:151) reaches the GPU (src/EffectiveAction.hpp:57).
//   * HipSU2MeasurementCorrelation -- the correlation measurement on the
//     device state (sp_measure_correlation), chosen by the factory in place of
//     SU2MeasurementCorrelation (src/FrgCoreFactory.cpp:30-46); it receives the
//     state object as the reference's measurements do
//     (src/Measurement.hpp:56) and reaches the context through it.
//   * HipSU2FrgCore -- computeStep / finalizeStep (src/FrgCore.hpp:77,86) and
//     the non-virtual takeMeasurements (src/FrgCore.hpp:36-68, non-deferred
//     branch).
//
// The interface classes below restate only the members of EffectiveAction,
// Measurement and FrgCore that the loop uses.  main() runs the loop of
// SpinParser::runCore (src/SpinParser.cpp:141-172) on a problem file written
// by the test suite and writes every correlation record; status codes become
// exceptions, which main turns into exit code 1 (src/SpinParser.cpp:88-92).
//
// Problem file (little endian): "SPP1", int32 nw, L, O, R, B, Rb, ncut; float32
// mesh[nw]; int32 site[L], inv[L], ovl_off[L+1], rid1[O], rid2[O],
// range_fwd[R], range_inv[R], corr[B*Rb]; float64 spin, cutoffs[ncut],
// v2[nw], v4ss[nv], v4dd[nv] (nv = nw^2 (nw+1)/2 * L).
// Output file: int32 nrec, int32 diverged; per record float64 cutoff,
// zz[B*Rb], dd[B*Rb].
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "spfrg.h"

namespace {

struct SpinParserError : std::runtime_error {
  int code;
  SpinParserError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void check(int rc, const sp_ctx *ctx, const char *what) {
  if (rc == SP_OK) return;
  throw SpinParserError(rc, std::string(what) + ": " + (ctx ? sp_last_error(ctx) : "error"));
}

// --- the interface members the loop uses -----------------------------------
struct EffectiveAction {
  virtual ~EffectiveAction() = default;
  virtual bool isDiverged() const = 0;
  float cutoff = 0.0f;
};

struct Measurement {
  Measurement(float minCutoff, float maxCutoff) : lo(minCutoff), hi(maxCutoff) {}
  virtual ~Measurement() = default;
  virtual void takeMeasurement(const EffectiveAction &state, bool isMasterTask) const = 0;
  float minCutoff() const { return lo; }
  float maxCutoff() const { return hi; }
  float lo, hi;
};

class FrgCore {
 public:
  virtual ~FrgCore() {
    for (Measurement *m : _measurements) delete m;
  }
  virtual void computeStep() = 0;
  virtual void finalizeStep(float newCutoff) = 0;
  void takeMeasurements() const {
    for (const Measurement *m : _measurements)
      if (_flowingFunctional->cutoff <= m->maxCutoff() && _flowingFunctional->cutoff >= m->minCutoff())
        m->takeMeasurement(*_flowingFunctional, true);
  }
  EffectiveAction *_flowingFunctional = nullptr;
  EffectiveAction *_flow = nullptr;
  std::vector<Measurement *> _measurements;

 protected:
  explicit FrgCore(const std::vector<Measurement *> &m) : _measurements(m) {}
};

// --- the GPU binding --------------------------------------------------------
class HipSU2EffectiveAction : public EffectiveAction {
 public:
  HipSU2EffectiveAction(sp_ctx *c, bool isFlow) : ctx(c), flow(isFlow) {}
  bool isDiverged() const override {
    if (!flow) return false;  // only the flow is checked (src/SpinParser.cpp:151)
    int d = 0;
    check(sp_is_diverged(ctx, &d), ctx, "sp_is_diverged");
    return d != 0;
  }
  sp_ctx *ctx;
  bool flow;
};

struct Record {
  double cutoff;
  std::vector<double> zz, dd;
};

class HipSU2MeasurementCorrelation : public Measurement {
 public:
  HipSU2MeasurementCorrelation(int n, std::vector<Record> *out)
      : Measurement(0.0f, 1e30f), nout(n), records(out) {}
  void takeMeasurement(const EffectiveAction &state, bool isMasterTask) const override {
    const auto &s = static_cast<const HipSU2EffectiveAction &>(state);
    Record r{state.cutoff, std::vector<double>(nout), std::vector<double>(nout)};
    check(sp_measure_correlation(s.ctx, r.zz.data(), r.dd.data()), s.ctx, "sp_measure_correlation");
    // duplicate cutoffs are dropped, as the reference's writer does
    if (isMasterTask && (records->empty() || (float)records->back().cutoff != (float)r.cutoff))
      records->push_back(std::move(r));
  }
  int nout;
  std::vector<Record> *records;
};

class HipSU2FrgCore : public FrgCore {
 public:
  HipSU2FrgCore(const sp_su2_problem &p, float cutoff0, const double *v2, const double *ss, const double *dd,
                const std::vector<Measurement *> &m)
      : FrgCore(m) {
    check(sp_ctx_create(&p, &ctx), ctx, "sp_ctx_create");
    check(sp_set_state(ctx, cutoff0, v2, ss, dd), ctx, "sp_set_state");
    _flowingFunctional = new HipSU2EffectiveAction(ctx, false);
    _flow = new HipSU2EffectiveAction(ctx, true);
    _flowingFunctional->cutoff = cutoff0;
    _flow->cutoff = cutoff0;
  }
  ~HipSU2FrgCore() override {
    delete _flowingFunctional;
    delete _flow;
    sp_ctx_destroy(ctx);
  }
  void computeStep() override {
    check(sp_compute_step(ctx), ctx, "sp_compute_step");
    _flow->cutoff = _flowingFunctional->cutoff;
  }
  void finalizeStep(float newCutoff) override {
    check(sp_finalize_step(ctx, newCutoff), ctx, "sp_finalize_step");
    _flowingFunctional->cutoff = newCutoff;
  }
  sp_ctx *ctx = nullptr;
};

// --- problem file -----------------------------------------------------------
struct Reader {
  std::ifstream in;
  explicit Reader(const char *path) : in(path, std::ios::binary) {
    if (!in) throw SpinParserError(SP_EARG, std::string("cannot open ") + path);
  }
  template <class T> T one() {
    T v;
    if (!in.read(reinterpret_cast<char *>(&v), sizeof(T))) throw SpinParserError(SP_EARG, "truncated problem file");
    return v;
  }
  template <class T> std::vector<T> many(size_t n) {
    std::vector<T> v(n);
    if (n && !in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T))))
      throw SpinParserError(SP_EARG, "truncated problem file");
    return v;
  }
};

int run(const char *problem_path, const char *out_path) {
  Reader r(problem_path);
  char magic[4];
  if (!r.in.read(magic, 4) || std::memcmp(magic, "SPP1", 4) != 0) throw SpinParserError(SP_EARG, "not a problem file");
  const int nw = r.one<int32_t>(), L = r.one<int32_t>(), O = r.one<int32_t>(), R = r.one<int32_t>();
  const int B = r.one<int32_t>(), Rb = r.one<int32_t>(), ncut = r.one<int32_t>();
  const auto mesh = r.many<float>(nw);
  const auto site = r.many<int32_t>(L), inv = r.many<int32_t>(L), off = r.many<int32_t>(L + 1);
  const auto rid1 = r.many<int32_t>(O), rid2 = r.many<int32_t>(O);
  const auto fwd = r.many<int32_t>(R), rinv = r.many<int32_t>(R), corr = r.many<int32_t>((size_t)B * Rb);
  const double spin = r.one<double>();
  const auto cutoffs = r.many<double>(ncut);
  const size_t nv = (size_t)nw * nw * (nw + 1) / 2 * L;
  const auto v2 = r.many<double>(nw), ss = r.many<double>(nv), dd = r.many<double>(nv);

  sp_su2_problem p{};
  p.nw = nw;
  p.mesh = mesh.data();
  p.L = L;
  p.site_rid = site.data();
  p.inv_rid = inv.data();
  p.ovl_off = off.data();
  p.ovl_rid1 = rid1.data();
  p.ovl_rid2 = rid2.data();
  p.R = R;
  p.range_fwd = fwd.data();
  p.range_inv = rinv.data();
  p.B = B;
  p.Rb = Rb;
  p.corr_rid = corr.data();
  p.spin = spin;
  p.precision = SP_PRECISION_FP64;

  std::vector<Record> records;
  std::vector<Measurement *> ms{new HipSU2MeasurementCorrelation(B * Rb, &records)};
  HipSU2FrgCore core(p, (float)cutoffs[0], v2.data(), ss.data(), dd.data(), ms);
  FrgCore *frg = &core;

  // SpinParser::runCore (src/SpinParser.cpp:141-172)
  bool diverged = false;
  int k = 0;
  while (k < ncut - 1) {
    frg->computeStep();
    frg->takeMeasurements();
    if (frg->_flow->isDiverged()) {
      diverged = true;
      break;
    }
    ++k;
    frg->finalizeStep((float)cutoffs[k]);
  }
  frg->takeMeasurements();

  std::ofstream out(out_path, std::ios::binary);
  const int32_t head[2] = {(int32_t)records.size(), diverged ? 1 : 0};
  out.write(reinterpret_cast<const char *>(head), sizeof head);
  for (const Record &rec : records) {
    out.write(reinterpret_cast<const char *>(&rec.cutoff), 8);
    out.write(reinterpret_cast<const char *>(rec.zz.data()), (std::streamsize)(rec.zz.size() * 8));
    out.write(reinterpret_cast<const char *>(rec.dd.data()), (std::streamsize)(rec.dd.size() * 8));
  }
  if (!out) throw SpinParserError(SP_EARG, std::string("cannot write ") + out_path);
  std::printf("runcore_su2: %zu records, diverged %d\n", records.size(), diverged ? 1 : 0);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: runcore_su2 PROBLEM_FILE OUTPUT_FILE\n");
    return 2;
  }
  try {
    return run(argv[1], argv[2]);
  } catch (const SpinParserError &e) {
    std::fprintf(stderr, "runcore_su2: error %d: %s\n", e.code, e.what());
    return 1;
  }
}
