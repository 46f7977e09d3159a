// Per-lane QUADPACK dqagse as a resumable state machine: every call to
// consume() takes one QK21 result, so a wavefront can run one QK21 for all of
// its lanes in lockstep while each lane advances its own adaptive integral
// (the bookkeeping, which differs per lane, is cheap next to the 21
// integrand evaluations).  The control flow and arithmetic follow dqagse
// exactly (Piessens et al. 1983), which gsl_integration_qags restates.
#pragma once

#include "upc_device.h"

namespace upc {

enum QagsStatus { kQagsRunning = 0, kQagsDone = 1, kQagsOverflow = 2 };

// CAP: interval-list capacity (1-based arrays).  The epsilon table holds at
// most one entry per bisection (numrl2 <= last + 1) plus the two qelg works
// past its end, and never more than dqelg's limexp = 50 (+2, 1-based).
template <int CAP>
struct Qags {
  static constexpr int NEPS = CAP + 6 < 55 ? CAP + 6 : 55;
  double alist[CAP + 1], blist[CAP + 1], rlist[CAP + 1], elist[CAP + 1];
  int iord[CAP + 1];
  double rlist2[NEPS], res3la[4];
  double a, b, epsabs, epsrel;
  double result, abserr, area, errsum, errbnd, errmax, erlarg, ertest, small, correc, defabs,
      erlast, a1, b1, a2, b2, area1, error1, defab1;
  int limit, ier, last, ierro, iroff1, iroff2, iroff3, ksgn, ktmin, maxerr, nrmax, nres, numrl2,
      extrap, noext, phase;

  __device__ void start(double a_, double b_, double epsabs_, double epsrel_, int limit_)
  {
    a = a_;
    b = b_;
    epsabs = epsabs_;
    epsrel = epsrel_;
    limit = limit_;
    phase = 0;
  }

  __device__ __forceinline__ void interval(double* qa, double* qb) const
  {
    if (phase == 0) {
      *qa = a;
      *qb = b;
    } else if (phase == 1) {
      *qa = a1;
      *qb = b1;
    } else {
      *qa = a2;
      *qb = b2;
    }
  }

  __device__ int begin_bisection()
  {
    if (last > CAP) return kQagsOverflow;
    a1 = alist[maxerr];
    b1 = 0.5 * (alist[maxerr] + blist[maxerr]);
    a2 = b1;
    b2 = blist[maxerr];
    erlast = errmax;
    phase = 1;
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
    for (int k = 1; k <= last; ++k) result = result + rlist[k];
    abserr = errsum;
    return finish_130();
  }
  __device__ int finish_130()
  {
    if (ier > 2) ier = ier - 1;
    return kQagsDone;
  }

  __device__ __forceinline__ int consume(const QK21Out& o)
  {
    constexpr double epmach = DBL_EPSILON, uflow = DBL_MIN, oflow = DBL_MAX;
    if (phase == 0) {
      result = o.result;
      abserr = o.abserr;
      defabs = o.resabs;
      const double resasc0 = o.resasc;
      const double dres = fabs(result);
      errbnd = fmax(epsabs, epsrel * dres);
      last = 1;
      alist[1] = a;
      blist[1] = b;
      rlist[1] = result;
      elist[1] = abserr;
      iord[1] = 1;
      ier = 0;
      if (abserr <= 100.0 * epmach * defabs && abserr > errbnd) ier = 2;
      if (limit == 1) ier = 1;
      if (ier != 0 || (abserr <= errbnd && abserr != resasc0) || abserr == 0.0) return kQagsDone;
      rlist2[1] = result;
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
    if (phase == 1) {
      area1 = o.result;
      error1 = o.abserr;
      defab1 = o.resasc;
      phase = 2;
      return kQagsRunning;
    }
    // phase 2: body of the dqagse main loop for this value of `last`
    const double area2 = o.result, error2 = o.abserr, defab2 = o.resasc;
    const double area12 = area1 + area2;
    const double erro12 = error1 + error2;
    errsum = errsum + erro12 - errmax;
    area = area + area12 - rlist[maxerr];
    if (!(defab1 == error1 || defab2 == error2)) {
      if (!(fabs(rlist[maxerr] - area12) > 1e-5 * fabs(area12) || erro12 < 0.99 * errmax)) {
        if (extrap) iroff2++;
        else iroff1++;
      }
      if (last > 10 && erro12 > errmax) iroff3++;
    }
    rlist[maxerr] = area1;
    rlist[last] = area2;
    errbnd = fmax(epsabs, epsrel * fabs(area));
    if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
    if (iroff2 >= 5) ierro = 3;
    if (last == limit) ier = 1;
    if (fmax(fabs(a1), fabs(b2)) <= (1.0 + 100.0 * epmach) * (fabs(a2) + 1000.0 * uflow)) ier = 4;
    if (error2 > error1) {
      alist[maxerr] = a2;
      alist[last] = a1;
      blist[last] = b1;
      rlist[maxerr] = area2;
      rlist[last] = area1;
      elist[maxerr] = error2;
      elist[last] = error1;
    } else {
      alist[last] = a2;
      blist[maxerr] = b1;
      blist[last] = b2;
      elist[maxerr] = error1;
      elist[last] = error2;
    }
    qpsrt(limit, last, &maxerr, &errmax, elist, iord, &nrmax);
    if (errsum <= errbnd) return finish_115();
    if (ier != 0) return finish_100();
    if (last == 2) {
      small = fabs(b - a) * 0.375;
      erlarg = errsum;
      ertest = errbnd;
      rlist2[2] = area;
      ++last;
      return begin_bisection();
    }
    if (noext) {
      ++last;
      return begin_bisection();
    }
    erlarg = erlarg - erlast;
    if (fabs(b1 - a1) > small) erlarg = erlarg + erro12;
    if (!extrap) {
      if (fabs(blist[maxerr] - alist[maxerr]) > small) {
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
        maxerr = iord[nrmax];
        errmax = elist[maxerr];
        if (fabs(blist[maxerr] - alist[maxerr]) > small) {
          ++last;
          return begin_bisection();
        }
        nrmax++;
      }
    }
    if (numrl2 + 3 > NEPS - 1) return kQagsOverflow;  // small tables only
    numrl2++;
    rlist2[numrl2] = area;
    double reseps, abseps;
    qelg(&numrl2, rlist2, &reseps, &abseps, res3la, &nres);
    ktmin++;
    if (ktmin > 5 && abserr < 1e-3 * errsum) ier = 5;
    if (abseps < abserr) {
      ktmin = 0;
      abserr = abseps;
      result = reseps;
      correc = erlarg;
      ertest = fmax(epsabs, epsrel * fabs(reseps));
      if (abserr <= ertest) return finish_100();
    }
    if (numrl2 == 1) noext = 1;
    if (ier == 5) return finish_100();
    maxerr = iord[1];
    errmax = elist[maxerr];
    nrmax = 1;
    extrap = 0;
    small = small * 0.5;
    erlarg = errsum;
    ++last;
    return begin_bisection();
  }

  __device__ int neval() const { return 42 * last - 21; }
};

}  // namespace upc
