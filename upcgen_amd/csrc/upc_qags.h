// Per-lane QUADPACK dqagse as a resumable state machine: every call to
// consume() takes the QK21 results of one step (the whole range first, then
// both halves of each bisection), so a wavefront can run the QK21 passes for
// all of its lanes in lockstep while each lane advances its own adaptive
// integral.
// The control flow and arithmetic follow dqagse exactly (Piessens et al.
// 1983), which gsl_integration_qags restates.
//
// Storage.  The scalar state lives in registers; only the lists that are
// indexed per lane (intervals, results, errors, the order list, the epsilon
// table) go through a store policy:
//  - QagsSmallStore<CAP> (fast path): per-lane private arrays, intervals
//    packed into 32 bits, the order list in 4-bit fields of one register;
//  - QagsBigStore (the reference's 1000-interval workspace, global memory).
// Every flux integral runs over q in [0, 10]; bisection keeps each interval a
// dyadic piece [10 m 2^-L, 10 (m + 1) 2^-L] of it, and those end points (and
// the midpoints 0.5 (a + b) that dqagse computes) are exact in FP64, so the
// packed form (L, m) expands to the same doubles the FP64 lists hold.
#pragma once

#include "upc_device.h"

namespace upc {

enum QagsStatus { kQagsRunning = 0, kQagsDone = 1, kQagsOverflow = 2 };

// ier of an integral stopped at its first QK21 pass with a NaN integrand
// value.  In the flux integrand NaN marks a node past the last form-factor
// knot, where the reference's gsl_spline_eval raises GSL_EDOM and its
// default error handler ends the program: the integral stops there rather
// than bisecting NaN values up to the 1000-interval limit.
constexpr int kQagsEdom = 100;

constexpr double kQagsA = 0.0, kQagsB = 10.0;  // integration range of every flux integral
constexpr double kQagsEps = 1e-4;               // epsabs = epsrel of the reference
constexpr int kQagsLimit = 1000;                // gsl_integration_qags limit of the reference

template <int CAP>
struct QagsSmallStore {
  static_assert(CAP >= 2 && CAP <= 16, "the order list packs 16 4-bit entries");
  static constexpr int cap = CAP;
  // The epsilon table holds at most one entry per bisection (numrl2 <= last
  // + 1) plus the two qelg works past its end (1-based).
  static constexpr int kEps = CAP + 6;
  using Iv = unsigned;  // L << 26 | m
  struct Arrays {
    Iv iv[CAP];
    double rl[CAP], el[CAP];
    double eps[kEps], r3[4];
  };
  Arrays* a;
  unsigned long long io;  // iord[k] - 1 in bits 4 (k - 1) .. 4 (k - 1) + 3
  int lim = CAP;          // intervals held before the integral is handed on (<= CAP)

  __device__ void reset() { io = 0ull; }
  __device__ int limit() const { return lim; }

  __device__ static Iv root() { return 0u; }
  __device__ static double scale(Iv v) { return ldexp(kQagsB, -(int)(v >> 26)); }
  __device__ static double lo(Iv v) { return (double)(v & 0x3ffffffu) * scale(v); }
  __device__ static double hi(Iv v) { return (double)((v & 0x3ffffffu) + 1u) * scale(v); }
  __device__ static double mid(Iv v) { return (double)(2u * (v & 0x3ffffffu) + 1u) * (0.5 * scale(v)); }
  __device__ static double width(Iv v) { return scale(v); }
  __device__ static Iv left(Iv v) { return (((v >> 26) + 1u) << 26) | (2u * (v & 0x3ffffffu)); }
  __device__ static Iv right(Iv v) { return left(v) | 1u; }

  __device__ Iv iv(int k) const { return a->iv[k - 1]; }
  __device__ void set_iv(int k, Iv v) { a->iv[k - 1] = v; }
  __device__ double rl(int k) const { return a->rl[k - 1]; }
  __device__ void set_rl(int k, double v) { a->rl[k - 1] = v; }
  __device__ double el(int k) const { return a->el[k - 1]; }
  __device__ void set_el(int k, double v) { a->el[k - 1] = v; }
  __device__ int io_at(int k) const { return (int)((io >> (4 * (k - 1))) & 15ull) + 1; }
  __device__ void set_io(int k, int v)
  {
    const int sh = 4 * (k - 1);
    io = (io & ~(15ull << sh)) | ((unsigned long long)(v - 1) << sh);
  }
  __device__ double* eps() { return a->eps; }
  __device__ double* r3() { return a->r3; }
};

// The reference's workspace: limit 1000, FP64 end points, one per lane in
// global memory.
struct QagsBigArrays {
  double alist[kQagsLimit], blist[kQagsLimit], rlist[kQagsLimit], elist[kQagsLimit];
  int iord[kQagsLimit];
  double eps[55], r3[4];  // dqelg's limexp = 50 (+2, 1-based)
};

struct QagsBigStore {
  static constexpr int cap = kQagsLimit;
  static constexpr int kEps = 55;
  using Iv = double2;  // (a, b)
  QagsBigArrays* a;

  __device__ void reset() {}
  __device__ static int limit() { return cap; }
  __device__ static Iv root() { return make_double2(kQagsA, kQagsB); }
  __device__ static double lo(Iv v) { return v.x; }
  __device__ static double hi(Iv v) { return v.y; }
  __device__ static double mid(Iv v) { return 0.5 * (v.x + v.y); }
  __device__ static double width(Iv v) { return fabs(v.y - v.x); }
  __device__ static Iv left(Iv v) { return make_double2(v.x, mid(v)); }
  __device__ static Iv right(Iv v) { return make_double2(mid(v), v.y); }

  __device__ Iv iv(int k) const { return make_double2(a->alist[k - 1], a->blist[k - 1]); }
  __device__ void set_iv(int k, Iv v)
  {
    a->alist[k - 1] = v.x;
    a->blist[k - 1] = v.y;
  }
  __device__ double rl(int k) const { return a->rlist[k - 1]; }
  __device__ void set_rl(int k, double v) { a->rlist[k - 1] = v; }
  __device__ double el(int k) const { return a->elist[k - 1]; }
  __device__ void set_el(int k, double v) { a->elist[k - 1] = v; }
  __device__ int io_at(int k) const { return a->iord[k - 1]; }
  __device__ void set_io(int k, int v) { a->iord[k - 1] = v; }
  __device__ double* eps() { return a->eps; }
  __device__ double* r3() { return a->r3; }
};

// dqpsrt over a store (1-based indices).
template <class S>
__device__ __forceinline__ void qpsrt(S& s, int last, int* maxerr, double* ermax, int* nrmax)
{
  constexpr int limit = kQagsLimit;
  int i, k, isucc, jupbn, jbnd, ibeg;
  double errmax, errmin;
  if (last <= 2) {
    s.set_io(1, 1);
    s.set_io(2, 2);
    goto L90;
  }
  errmax = s.el(*maxerr);
  if (*nrmax != 1) {
    const int ido = *nrmax - 1;
    for (i = 1; i <= ido; ++i) {
      isucc = s.io_at(*nrmax - 1);
      if (errmax <= s.el(isucc)) break;
      s.set_io(*nrmax, isucc);
      (*nrmax)--;
    }
  }
  jupbn = last;
  if (last > (limit / 2 + 2)) jupbn = limit + 3 - last;
  errmin = s.el(last);
  jbnd = jupbn - 1;
  ibeg = *nrmax + 1;
  if (ibeg <= jbnd) {
    for (i = ibeg; i <= jbnd; ++i) {
      isucc = s.io_at(i);
      if (errmax >= s.el(isucc)) goto L60;
      s.set_io(i - 1, isucc);
    }
  }
  s.set_io(jbnd, *maxerr);
  s.set_io(jupbn, last);
  goto L90;
L60:
  s.set_io(i - 1, *maxerr);
  k = jbnd;
  for (int j = i; j <= jbnd; ++j) {
    isucc = s.io_at(k);
    if (errmin < s.el(isucc)) {
      s.set_io(k + 1, last);
      goto L90;
    }
    s.set_io(k + 1, isucc);
    k--;
  }
  s.set_io(i, last);
L90:
  *maxerr = s.io_at(*nrmax);
  *ermax = s.el(*maxerr);
}

template <class S>
struct Qags {
  using Iv = typename S::Iv;
  static constexpr int limit = kQagsLimit;
  S s;
  Iv cur;  // the interval being bisected (the root before the first QK21)
  double result, abserr, area, errsum, errbnd, errmax, erlarg, ertest, small, correc, defabs, erlast;
  int ier, last, ierro, iroff1, iroff2, iroff3, ksgn, ktmin, maxerr, nrmax, nres, numrl2, extrap, noext, phase;

  __device__ void start()
  {
    s.reset();
    cur = S::root();
    phase = 0;
  }

  // QK21 passes of the next step: phase 0 one (the whole range), phase 2
  // two (k = 0: left half of `cur`, k = 1: right half).
  __device__ __forceinline__ int passes() const { return phase == 0 ? 1 : 2; }
  __device__ __forceinline__ void interval(int k, double* qa, double* qb) const
  {
    *qa = phase != 0 && k == 1 ? S::mid(cur) : S::lo(cur);
    *qb = phase != 0 && k == 0 ? S::mid(cur) : S::hi(cur);
  }

  __device__ int begin_bisection()
  {
    if (last > s.limit()) return kQagsOverflow;
    cur = s.iv(maxerr);
    erlast = errmax;
    phase = 2;
    return kQagsRunning;
  }

  // Tail of dqagse after the main loop (labels 100..130).
  __device__ int finish_100()
  {
    if (abserr == DBL_MAX) return finish_115();
    if (ier + ierro == 0) return finish_110();
    if (ierro == 3) abserr = abserr + correc;
    if (ier == 0) ier = 3;
    if (result != 0.0 && area != 0.0) {
      if (abserr / fabs(result) > errsum / fabs(area)) return finish_115();
      return finish_110();
    }
    if (abserr > errsum) return finish_115();
    if (area == 0.0) return finish_130();
    return finish_110();
  }
  __device__ int finish_110()
  {
    if (ksgn == -1 && fmax(fabs(result), fabs(area)) <= defabs * 0.01) return finish_130();
    if (0.01 > (result / area) || (result / area) > 100.0 || errsum > fabs(area)) ier = 6;
    return finish_130();
  }
  __device__ int finish_115()
  {
    result = 0.0;
    for (int k = 1; k <= last; ++k) result = result + s.rl(k);
    abserr = errsum;
    return finish_130();
  }
  __device__ int finish_130()
  {
    if (ier > 2) ier = ier - 1;
    return kQagsDone;
  }

  // o: the pass on the whole range (phase 0) or the left half; o2: the right
  // half (phase 2 only).
  __device__ __forceinline__ int consume(const QK21Out& o, const QK21Out& o2)
  {
    constexpr double epmach = DBL_EPSILON, uflow = DBL_MIN, oflow = DBL_MAX;
    if (phase == 0) {
      result = o.result;
      abserr = o.abserr;
      defabs = o.resabs;
      last = 1;
      if (result != result) {
        ier = kQagsEdom;
        return kQagsDone;
      }
      const double resasc0 = o.resasc;
      const double dres = fabs(result);
      errbnd = fmax(kQagsEps, kQagsEps * dres);
      s.set_iv(1, cur);
      s.set_rl(1, result);
      s.set_el(1, abserr);
      s.set_io(1, 1);
      ier = 0;
      if (abserr <= 100.0 * epmach * defabs && abserr > errbnd) ier = 2;
      if (limit == 1) ier = 1;
      if (ier != 0 || (abserr <= errbnd && abserr != resasc0) || abserr == 0.0) return kQagsDone;
      s.eps()[1] = result;
      errmax = abserr;
      maxerr = 1;
      area = result;
      errsum = abserr;
      abserr = oflow;
      nrmax = 1;
      nres = 0;
      numrl2 = 2;
      ktmin = 0;
      extrap = 0;
      noext = 0;
      ierro = 0;
      iroff1 = iroff2 = iroff3 = 0;
      erlarg = ertest = small = correc = 0.0;
      ksgn = -1;
      if (dres >= (1.0 - 50.0 * epmach) * defabs) ksgn = 1;
      last = 2;
      return begin_bisection();
    }
    // phase 2: body of the dqagse main loop for this value of `last`
    const double a1 = S::lo(cur), b1 = S::mid(cur), a2 = b1, b2 = S::hi(cur);
    const double area1 = o.result, error1 = o.abserr, defab1 = o.resasc;
    const double area2 = o2.result, error2 = o2.abserr, defab2 = o2.resasc;
    const double area12 = area1 + area2;
    const double erro12 = error1 + error2;
    if (area12 != area12) {
      result = area12;
      ier = kQagsEdom;
      return kQagsDone;
    }
    const double rmax = s.rl(maxerr);
    errsum = errsum + erro12 - errmax;
    area = area + area12 - rmax;
    if (!(defab1 == error1 || defab2 == error2)) {
      if (!(fabs(rmax - area12) > 1e-5 * fabs(area12) || erro12 < 0.99 * errmax)) {
        if (extrap) iroff2++;
        else iroff1++;
      }
      if (last > 10 && erro12 > errmax) iroff3++;
    }
    errbnd = fmax(kQagsEps, kQagsEps * fabs(area));
    if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
    if (iroff2 >= 5) ierro = 3;
    if (last == limit) ier = 1;
    if (fmax(fabs(a1), fabs(b2)) <= (1.0 + 100.0 * epmach) * (fabs(a2) + 1000.0 * uflow)) ier = 4;
    // dqagse stores (a1, b1, area1, error1) at maxerr and (a2, b2, area2,
    // error2) at last, then swaps them when error2 > error1: the final
    // contents are written once here.
    const bool swap = error2 > error1;
    s.set_iv(maxerr, swap ? S::right(cur) : S::left(cur));
    s.set_iv(last, swap ? S::left(cur) : S::right(cur));
    s.set_rl(maxerr, swap ? area2 : area1);
    s.set_rl(last, swap ? area1 : area2);
    s.set_el(maxerr, swap ? error2 : error1);
    s.set_el(last, swap ? error1 : error2);
    qpsrt(s, last, &maxerr, &errmax, &nrmax);
    if (errsum <= errbnd) return finish_115();
    if (ier != 0) return finish_100();
    if (last == 2) {
      small = fabs(kQagsB - kQagsA) * 0.375;
      erlarg = errsum;
      ertest = errbnd;
      s.eps()[2] = area;
      ++last;
      return begin_bisection();
    }
    if (noext) {
      ++last;
      return begin_bisection();
    }
    erlarg = erlarg - erlast;
    if (S::width(S::left(cur)) > small) erlarg = erlarg + erro12;
    if (!extrap) {
      if (S::width(s.iv(maxerr)) > small) {
        ++last;
        return begin_bisection();
      }
      extrap = 1;
      nrmax = 2;
    }
    if (!(ierro == 3 || erlarg <= ertest)) {
      const int id = nrmax;
      int jupbnd = last;
      if (last > (2 + limit / 2)) jupbnd = limit + 3 - last;
      for (int k = id; k <= jupbnd; ++k) {
        maxerr = s.io_at(nrmax);
        errmax = s.el(maxerr);
        if (S::width(s.iv(maxerr)) > small) {
          ++last;
          return begin_bisection();
        }
        nrmax++;
      }
    }
    if (numrl2 + 3 > S::kEps - 1) return kQagsOverflow;  // small tables only
    numrl2++;
    s.eps()[numrl2] = area;
    double reseps, abseps;
    qelg(&numrl2, s.eps(), &reseps, &abseps, s.r3(), &nres);
    ktmin++;
    if (ktmin > 5 && abserr < 1e-3 * errsum) ier = 5;
    if (abseps < abserr) {
      ktmin = 0;
      abserr = abseps;
      result = reseps;
      correc = erlarg;
      ertest = fmax(kQagsEps, kQagsEps * fabs(reseps));
      if (abserr <= ertest) return finish_100();
    }
    if (numrl2 == 1) noext = 1;
    if (ier == 5) return finish_100();
    maxerr = s.io_at(1);
    errmax = s.el(maxerr);
    nrmax = 1;
    extrap = 0;
    small = small * 0.5;
    erlarg = errsum;
    ++last;
    return begin_bisection();
  }

  __device__ int neval() const { return 42 * last - 21; }

  // Status of a finished integral: dqagse's ier, or kQagsEdom.
  __device__ int status() const { return ier; }
};

}  // namespace upc
