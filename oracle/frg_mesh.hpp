// Frequency mesh and trapezoid integrators of the CPU oracle -- TEST
// INFRASTRUCTURE ONLY (the checker of the HIP path; nothing in the product
// links it).
//
// Reference files followed (paths relative to the reference repository):
//   src/FrequencyDiscretization.hpp:253-356   lesser / greater / offset / interpolateOffset
//   src/lib/Integrator.hpp:137-287            ImplicitIntegrator (trapezoid, off-mesh ends)
//
// The integrators are templated on the bundle type Buf, which provides
// reset(), add(const Buf&), mult_add(T, const Buf&) and scale(T).
#pragma once
#include <cmath>
#include <vector>

namespace frg {

template <class T> struct Mesh {
  int n;
  std::vector<T> v;  // 2n values, v[n + i] = mesh point i, i in [-n, n)
  Mesh(int nw, const float *mesh) : n(nw), v(2 * nw) {
    for (int i = 0; i < n; ++i) {
      v[n + i] = T(mesh[i]);
      v[n - 1 - i] = -T(mesh[i]);
    }
  }
  T at(int i) const { return v[n + i]; }
  int lesser(T w) const {
    if (w < 0) return -(greater(-w) + 1);
    if (w <= at(0)) return 0;
    for (int i = 1; i < n; ++i) if (at(i) > w) return i - 1;
    return n - 1;
  }
  int greater(T w) const {
    if (w < 0) return -(lesser(-w) + 1);
    if (w <= at(0)) return 0;
    for (int i = 1; i < n; ++i) if (at(i) > w) return i;
    return n - 1;
  }
  int offset(T w) const {
    if (w <= at(0)) return 0;
    for (int i = 1; i < n; ++i) if (at(i) >= w) return i;
    return n - 1;
  }
  void interp(T w, int &lo, int &hi, T &bias) const {
    if (w <= at(0)) { lo = 0; hi = 0; bias = T(0); return; }
    for (int i = 1; i < n; ++i) {
      if (at(i) > w) {
        hi = i; lo = i - 1;
        bias = (w - at(lo)) / (at(hi) - at(lo));
        return;
      }
    }
    lo = n - 1; hi = n - 1; bias = T(0);
  }
};

// ImplicitIntegrator::integrateWithObscureLeftBoundary: [min, mesh point max]
template <class T, class Buf, class F>
void integrate_left(const Mesh<T> &m, T min, int max, F &&f, Buf &ib, Buf &rb, long &calls) {
  rb.reset();
  int umin = m.greater(min);
  if (umin != max) {
    f(min, ib); ++calls; rb.mult_add(m.at(umin) - min, ib);
    f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - min, ib);
    while (++umin != max) { f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin - 1), ib); }
    f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin) - m.at(umin - 1), ib);
    rb.scale(T(0.5));
  } else {
    f(min, ib); ++calls; rb.add(ib);
    f(m.at(umin), ib); ++calls; rb.add(ib);
    rb.scale(T(0.5) * (m.at(umin) - min));
  }
}

// integrateWithObscureRightBoundary: [mesh point min, max]
template <class T, class Buf, class F>
void integrate_right(const Mesh<T> &m, int min, T max, F &&f, Buf &ib, Buf &rb, long &calls) {
  rb.reset();
  int umin = min;
  int umax = m.lesser(max);
  if (umin != umax) {
    f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin), ib);
    while (++umin != umax) { f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin - 1), ib); }
    f(m.at(umin), ib); ++calls; rb.mult_add(max - m.at(umin - 1), ib);
    f(max, ib); ++calls; rb.mult_add(max - m.at(umin), ib);
    rb.scale(T(0.5));
  } else {
    f(max, ib); ++calls; rb.add(ib);
    f(m.at(umin), ib); ++calls; rb.add(ib);
    rb.scale(T(0.5) * (max - m.at(umin)));
  }
}

// integrateWithObscureBoundaries: [min, max], both off the mesh
template <class T, class Buf, class F>
void integrate_both(const Mesh<T> &m, T min, T max, F &&f, Buf &ib, Buf &rb, long &calls) {
  rb.reset();
  int umin = m.greater(min);
  int umax = m.lesser(max);
  if (umax >= umin) {
    f(min, ib); ++calls; rb.mult_add(m.at(umin) - min, ib);
    if (umax != umin) {
      f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - min, ib);
      while (++umin != umax) { f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin - 1), ib); }
      f(m.at(umin), ib); ++calls; rb.mult_add(max - m.at(umin - 1), ib);
    } else {
      f(m.at(umin), ib); ++calls; rb.mult_add(max - min, ib);
    }
    f(max, ib); ++calls; rb.mult_add(max - m.at(umin), ib);
    rb.scale(T(0.5));
  } else {
    f(max, ib); ++calls; rb.add(ib);
    f(min, ib); ++calls; rb.add(ib);
    rb.scale(T(0.5) * (max - min));
  }
}

// The three integration windows of one channel (src/SU2/SU2FrgCore.cpp:373-424,
// src/TRI/TRIFrgCore.cpp:2943-2993): (-inf, -(X+cut)], [cut-X, -cut], [cut, inf)
// clipped to the mesh; each window's result is added to cur.
template <class T, class Buf, class F>
void katanin_windows(const Mesh<T> &m, T X, T cutoff, F &&f, Buf &ib, Buf &rb, Buf &cur, long &calls) {
  const int nw = m.n;
  const T wmin = m.at(-nw), wmax = m.at(nw - 1);
  if (-(X + cutoff) > wmin) { integrate_right(m, -nw, -(X + cutoff), f, ib, rb, calls); cur.add(rb); }
  if (X - cutoff > cutoff) { integrate_both(m, cutoff - X, -cutoff, f, ib, rb, calls); cur.add(rb); }
  if (cutoff < wmax) { integrate_left(m, cutoff, nw - 1, f, ib, rb, calls); cur.add(rb); }
}

template <class T> T two_pi() { return T(2.0) * T(M_PI); }
template <> inline float two_pi<float>() { return 2.0f * (float)M_PI; }

}  // namespace frg
