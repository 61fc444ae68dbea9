// Known-answer harness over the reference sources -- TEST INFRASTRUCTURE.
//
// Built by oracle/Makefile.ref from the unmodified reference translation units
// (outputs under oracle/_ref/ only).  It reads commands on stdin and prints the
// reference's results, which tests/golden/make_ref_fixtures.py stores as
// fixtures.  The reference's lautil.cpp includes mkl.h, which this image lacks,
// so it is not compiled: nothing here reaches CSolExpCenter::solveFH.
// Access specifiers are opened by the build flags of this file only, to read
// per-sphere constants that the reference keeps private.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include <new>

#include "BDnam.h"
#include "readutil.h"
#include "contact.h"
#include "expansion.h"
#include "molecule.h"
#include "rotcoeff.h"
#include "system.h"
#include "transcoeff.h"
#include "triexpan.h"
#include "xforms.h"

static double rd() {
  double v;
  if (scanf("%lf", &v) != 1) exit(3);
  return v;
}
static int ri() { return (int)rd(); }
static void out(const std::vector<double> &v) {
  for (size_t i = 0; i < v.size(); i++) printf("%.17g ", v[i]);
  printf("\n");
  fflush(stdout);
}
static std::vector<double> rdv(int n) {
  std::vector<double> v(n);
  for (int i = 0; i < n; i++) v[i] = rd();
  return v;
}
static std::vector<double> padded(const CExpan &E, int p) {
  std::vector<double> v(p * p, 0.0);
  const std::vector<double> &M = E.getVector();
  for (size_t i = 0; i < M.size() && (int)i < p * p; i++) v[i] = M[i];
  return v;
}

// One molecule built with the constructor CBDnam uses (molecule.cpp:422-487).  Reads:
// ns (x y z r)*ns nq (x y z q)*nq [F(ns*N^2) H(ns*N^2) with withFH], centres relative to the
// molecule centre.  Without withFH the self expansions are zero (makeMove never reads them).
static CMolecule *bd_molecule(int moltype, bool withFH = false) {
  int ns = ri();
  std::vector<CPnt> cens(ns);
  std::vector<double> rad(ns);
  for (int k = 0; k < ns; k++) {
    double x = rd(), y = rd(), z = rd();
    cens[k] = CPnt(x, y, z);
    rad[k] = rd();
  }
  int nq = ri();
  std::vector<CPnt> pos(nq);
  std::vector<double> q(nq);
  for (int i = 0; i < nq; i++) {
    double x = rd(), y = rd(), z = rd();
    pos[i] = CPnt(x, y, z);
    q[i] = rd();
  }
  const double idiel = 4.0, intraRc = 30.0;
  int L = N_POLES * N_POLES;
  std::vector<double> zero(L, 0.0);
  std::vector<double> fh = withFH ? rdv(2 * ns * L) : std::vector<double>(2 * ns * L, 0.0);
  std::vector<CMulExpan *> Fs(ns), Hs(ns);
  for (int k = 0; k < ns; k++) {
    Fs[k] = new CMulExpan(&fh[k * L], CRange(0, N_POLES), rad[k]);
    Hs[k] = new CMulExpan(&fh[(ns + k) * L], CRange(0, N_POLES), rad[k]);
  }
  std::vector<std::vector<CPnt> > SPxes;
  std::vector<int> nSPx;
  std::vector<std::vector<int> > neighs, nearL, farL;
  CMolecule::generateMolSPX(cens, rad, SPxes, nSPx, neighs);
  CMolecule::generateMolTypeIntraPolLists(cens, rad, intraRc, nearL, farL);
  std::vector<std::vector<double> > qF, qH;
  std::vector<double> tF, tH;
  CMolecule::generateMolExposedChargesFH(Fs, Hs, cens, rad, SPxes, nSPx, qF, qH, tF, tH);
  std::vector<CLocalExpan *> LF, LH, LFf, LHf;
  CMolecule::computeMolTypeValues(CPnt(), cens, rad, q, pos, idiel, intraRc, SPxes, nSPx, neighs, Fs, Hs,
                                  qF, qH, tF, tH, nearL, farL, LF, LH, LFf, LHf);
  std::vector<REAL *> iMats(ns, (REAL *)0);
  std::vector<CMolCell> cells;
  return new CMolecule(moltype, CPnt(), cens, rad, q, pos, idiel, iMats, intraRc, SPxes, nSPx, neighs, nearL,
                       Fs, Hs, LF, LH, LFf, LHf, qF, qH, tF, tH, cells);
}

int main() {
  std::cout.precision(17);
  char cmd[64];
  bool inited = false;
  while (scanf("%63s", cmd) == 1) {
    std::string c(cmd);
    if (c == "init") {  // kappa sdiel
      double kappa = rd(), sdiel = rd();
      // silence the reference's progress printing while it sets up its tables
      std::streambuf *old = std::cout.rdbuf(0);
      CSystem::initConstants(kappa, sdiel, 298.15);
      std::cout.rdbuf(old);
      inited = true;
      printf("ok %d\n", N_POLES);
      fflush(stdout);
      continue;
    }
    if (!inited) return 2;
    std::streambuf *old = std::cout.rdbuf(0);  // reference prints diagnostics to cout
    std::vector<double> res;
    if (c == "sh") {  // P theta phi
      int P = ri();
      double th = rd(), ph = rd();
      CSHExpan Y(th, ph, P, true);
      res = padded(Y, P);
    } else if (c == "besk" || c == "besi") {  // L val
      int L = ri();
      double val = rd();
      std::vector<double> K(L, 1.0);
      if (c == "besk") CLExpan::BESSEL(K, val);
      else CMExpan::BESSEL(K, val);
      res = K;
    } else if (c == "lexp" || c == "mexp") {  // p bkappa scale n (q x y z)*n
      int p = ri();
      bool bk = ri() != 0;
      double scale = rd();
      int n = ri();
      std::vector<double> q(n);
      std::vector<CPnt> pos(n);
      for (int i = 0; i < n; i++) {
        q[i] = rd();
        double x = rd(), y = rd(), z = rd();
        pos[i] = CPnt(x, y, z);
      }
      if (c == "lexp") res = padded(CLocalExpan(q, pos, p, bk, scale), p);
      else res = padded(CMulExpan(q, pos, p, bk, scale), p);
    } else if (c == "gradlocal") {  // p bkappa scale n (gx gy gz x y z)*n
      int p = ri();
      bool bk = ri() != 0;
      double scale = rd();
      int n = ri();
      std::vector<CPnt> g(n), pos(n);
      for (int i = 0; i < n; i++) {
        double a = rd(), b = rd(), d = rd();
        g[i] = CPnt(a, b, d);
        double x = rd(), y = rd(), z = rd();
        pos[i] = CPnt(x, y, z);
      }
      CGradExpan G(&g[0], pos, p, bk, scale, false);
      for (int d = 0; d < 3; d++) {
        std::vector<double> v = padded(G[d], p);
        res.insert(res.end(), v.begin(), v.end());
      }
    } else if (c == "rotate") {  // theta phi xi p p2 fwd block(p2^2)
      double th = rd(), ph = rd(), xi = rd();
      int p = ri(), p2 = ri(), fwd = ri();
      std::vector<double> v = rdv(p2 * p2);
      CRotCoeff R(false);
      R.reset(th, ph, xi, p);
      CMulExpan Min(&v[0], CRange(0, p2), 1.0), Mout;
      R.rotate(Min, Mout, 1, p2, fwd != 0);
      res = padded(Mout, p2);
      res.push_back(R.isSingular() ? 1.0 : 0.0);
    } else if (c == "drot") {  // theta phi xi p which fwd block(p^2)
      double th = rd(), ph = rd(), xi = rd();
      int p = ri(), which = ri(), fwd = ri();
      std::vector<double> v = rdv(p * p);
      CRotCoeff R(true);
      R.reset(th, ph, xi, p);
      CMulExpan Min(&v[0], CRange(0, p), 1.0), Mout;
      if (which == 0) R.dRotateT(Min, Mout, fwd != 0);
      else R.dRotateP(Min, Mout, fwd != 0);
      res = padded(Mout, p);
    } else if (c == "rotquat") {  // w x y z p fwd block(p^2)
      double w = rd(), x = rd(), y = rd(), z = rd();
      int p = ri(), fwd = ri();
      std::vector<double> v = rdv(p * p);
      CQuat Q(w, CPnt(x, y, z));
      CRotCoeff R(false);
      R.reset(Q, p);
      CMulExpan Min(&v[0], CRange(0, p), 1.0), Mout;
      R.rotateWithXi(Min, Mout, 1, p, fwd != 0);
      res = padded(Mout, p);
    } else if (c == "trans") {  // rho kappa p scalei scaleo grad block(p^2)
      double rho = rd(), kap = rd();
      int p = ri();
      double si = rd(), so = rd();
      int grad = ri();
      std::vector<double> v = rdv(p * p);
      CTransCoeff T(grad != 0);
      T.initScale(si, so);
      T.reset(rho, kap, p);
      CMulExpan Min(&v[0], CRange(0, p), si);  // stores scale 1/si as the apply expects
      CLocalExpan Mout;
      if (grad) T.dTranslate(Min, Mout, false);
      else T.translate(Min, Mout, false);
      res = padded(Mout, p);
    } else if (c == "xformh" || c == "xformhg") {  // Px Py Pz asrc atgt p block(p^2)
      double P[3] = {rd(), rd(), rd()};
      double as = rd(), at = rd();
      int p = ri();
      std::vector<double> v = rdv(p * p);
      CMulExpan Hs(&v[0], CRange(0, p), as);
      CSolExpCenter src(0, CPnt(), as, &Hs), tgt(1, CPnt(), at, &Hs);
      bool g = (c == "xformhg");
      CXFormA X(src, tgt, g);
      X.reset(CPnt(P[0], P[1], P[2]), p);
      if (!g) {
        CLocalExpan L;
        X.xformH(Hs, L, true);
        res = padded(L, p);
      } else {
        CGradExpan G;
        X.xformH(Hs, G, true);
        for (int d = 0; d < 3; d++) {
          std::vector<double> b = padded(G[d], p);
          res.insert(res.end(), b.begin(), b.end());
        }
      }
    } else if (c == "xformi") {  // Px Py Pz asrc atgt pH pF H(p^2) F(p^2), p = max
      double P[3] = {rd(), rd(), rd()};
      double as = rd(), at = rd();
      int pH = ri(), pF = ri();
      int p = pH > pF ? pH : pF;
      std::vector<double> h = rdv(p * p), f = rdv(p * p);
      CMulExpan Hs(&h[0], CRange(0, p), as), Fs(&f[0], CRange(0, p), as);
      CSolExpCenter src(0, CPnt(), as, &Hs), tgt(1, CPnt(), at, &Hs);
      CXFormAIntra X(src, tgt);
      X.reset(CPnt(P[0], P[1], P[2]), pH, pF);
      CLocalExpan LF, LH;
      X.xformF(Fs, LF, true);
      X.xformH(Hs, LH, true);
      res = padded(LH, p);
      std::vector<double> b = padded(LF, p);
      res.insert(res.end(), b.begin(), b.end());
    } else if (c == "xformn") {  // Px Py Pz asrc atgt tQF tQH n (x y z qF qH)*n ; order N_POLES
      double P[3] = {rd(), rd(), rd()};
      double as = rd(), at = rd(), tQF = rd(), tQH = rd();
      int n = ri();
      std::vector<CPnt> spx(n);
      std::vector<double> qF(n), qH(n);
      for (int i = 0; i < n; i++) {
        double x = rd(), y = rd(), z = rd();
        spx[i] = CPnt(x, y, z);
        qF[i] = rd();
        qH[i] = rd();
      }
      CMulExpan Hs(CRange(0, N_POLES), as);
      CSolExpCenter src(0, CPnt(), as, &Hs, spx, n), tgt(1, CPnt(), at, &Hs);
      CLocalExpan LF, LH;
      CXFormNIntra::xformFH(src, tgt, LF, LH, CPnt(P[0], P[1], P[2]), &qF[0], &qH[0], tQF, tQH);
      res = padded(LH, N_POLES);
      std::vector<double> b = padded(LF, N_POLES);
      res.insert(res.end(), b.begin(), b.end());
    } else if (c == "order") {  // rho tq scale
      double rho = rd(), tq = rd(), sc = rd();
      res.push_back(CXFormBase::computeOrder(rho, tq, sc));
    } else if (c == "spx") {  // N r  (reference spiral scaled to radius r)
      int N = ri();
      double r = rd();
      std::vector<REAL> th, ph;
      CMolecule::spherePts(N, th, ph);
      for (int i = 0; i < N; i++) {
        CPnt q = SphToCart(CSpPnt(1, th[i], ph[i]));
        CPnt pt = r * q;
        res.push_back(pt.x());
        res.push_back(pt.y());
        res.push_back(pt.z());
      }
    } else if (c == "imat") {  // npe npb (xyz)*npe (xyz)*npb  -> N_POLES^4 column-major
      int npe = ri(), npb = ri();
      std::vector<CPnt> spe(npe), spb(npb);
      for (int i = 0; i < npe; i++) { double x = rd(), y = rd(), z = rd(); spe[i] = CPnt(x, y, z); }
      for (int i = 0; i < npb; i++) { double x = rd(), y = rd(), z = rd(); spb[i] = CPnt(x, y, z); }
      std::vector<double> M(N_POLES * N_POLES * N_POLES * N_POLES);
      CSolExpCenter::build_IMat(spe, spb, &M[0]);
      res = M;
    } else if (c == "moltype") {
      // ns (x y z r)*ns nq (x y z q)*nq rcen(3) idiel intraRc
      //   -> per sphere: nexp nspx neighbours-count near-count [spx xyz]
      //      consts(10*N) Dfix Vfix E0 qH(init)
      int ns = ri();
      std::vector<CPnt> cens(ns);
      std::vector<double> rad(ns);
      for (int k = 0; k < ns; k++) {
        double x = rd(), y = rd(), z = rd();
        cens[k] = CPnt(x, y, z);
        rad[k] = rd();
      }
      int nq = ri();
      std::vector<CPnt> pos(nq);
      std::vector<double> q(nq);
      for (int i = 0; i < nq; i++) {
        double x = rd(), y = rd(), z = rd();
        pos[i] = CPnt(x, y, z);
        q[i] = rd();
      }
      double rc[3] = {rd(), rd(), rd()};
      double idiel = rd(), intraRc = rd();
      std::vector<std::vector<CPnt> > SPxes;
      std::vector<int> nSPx;
      std::vector<std::vector<int> > neighs, nearL, farL;
      CMolecule::generateMolSPX(cens, rad, SPxes, nSPx, neighs);
      CMolecule::generateMolTypeIntraPolLists(cens, rad, intraRc, nearL, farL);
      std::vector<REAL *> iMats(ns, (REAL *)0);
      CMolecule mol(CPnt(rc[0], rc[1], rc[2]), cens, rad, q, pos, idiel, iMats, intraRc, SPxes, nSPx,
                    neighs, nearL);
      for (int k = 0; k < ns; k++) {
        CSolExpCenter &S = mol.getKS(k);
        res.push_back(S.getSPExSize());
        res.push_back(nSPx[k]);
        res.push_back(neighs[k].size());
        res.push_back(nearL[k].size());
        for (size_t h = 0; h < SPxes[k].size(); h++) {
          res.push_back(SPxes[k][h].x());
          res.push_back(SPxes[k][h].y());
          res.push_back(SPxes[k][h].z());
        }
        double *cs[10] = {S.m_constH1, S.m_constF1, S.m_constLH1, S.m_constLF1, S.m_constH2,
                          S.m_constF2, S.m_constLH2, S.m_constLF2, S.m_constInvH, S.m_constChgH};
        for (int a = 0; a < 10; a++)
          for (int n = 0; n < N_POLES; n++) res.push_back(cs[a][n]);
        for (int i = 0; i < N_POLES * N_POLES; i++) res.push_back(S.m_Dfix[i]);
        for (int i = 0; i < N_POLES * N_POLES; i++) res.push_back(S.m_Vfix[i]);
        std::vector<double> e = padded(S.getH(), N_POLES);
        res.insert(res.end(), e.begin(), e.end());
        for (int h = 0; h < S.getSPExSize(); h++) res.push_back(S.m_qSolvedH[h]);
      }
    } else if (c == "molvals") {
      // ns (x y z r)*ns nq (x y z q)*nq rcen(3) idiel intraRc F(ns*N^2) H(ns*N^2)
      //   -> per sphere qF qH totals, then LF LH LFfar LHfar
      int ns = ri();
      std::vector<CPnt> cens(ns);
      std::vector<double> rad(ns);
      for (int k = 0; k < ns; k++) {
        double x = rd(), y = rd(), z = rd();
        cens[k] = CPnt(x, y, z);
        rad[k] = rd();
      }
      int nq = ri();
      std::vector<CPnt> pos(nq);
      std::vector<double> q(nq);
      for (int i = 0; i < nq; i++) {
        double x = rd(), y = rd(), z = rd();
        pos[i] = CPnt(x, y, z);
        q[i] = rd();
      }
      double rc[3] = {rd(), rd(), rd()};
      double idiel = rd(), intraRc = rd();
      int L = N_POLES * N_POLES;
      std::vector<CMulExpan *> Fs(ns), Hs(ns);
      for (int k = 0; k < ns; k++) {
        std::vector<double> f = rdv(L);
        Fs[k] = new CMulExpan(&f[0], CRange(0, N_POLES), rad[k]);
      }
      for (int k = 0; k < ns; k++) {
        std::vector<double> h = rdv(L);
        Hs[k] = new CMulExpan(&h[0], CRange(0, N_POLES), rad[k]);
      }
      std::vector<std::vector<CPnt> > SPxes;
      std::vector<int> nSPx;
      std::vector<std::vector<int> > neighs, nearL, farL;
      CMolecule::generateMolSPX(cens, rad, SPxes, nSPx, neighs);
      CMolecule::generateMolTypeIntraPolLists(cens, rad, intraRc, nearL, farL);
      std::vector<std::vector<double> > qF, qH;
      std::vector<double> tF, tH;
      CMolecule::generateMolExposedChargesFH(Fs, Hs, cens, rad, SPxes, nSPx, qF, qH, tF, tH);
      std::vector<CLocalExpan *> LF, LH, LFf, LHf;
      CMolecule::computeMolTypeValues(CPnt(rc[0], rc[1], rc[2]), cens, rad, q, pos, idiel, intraRc, SPxes,
                                      nSPx, neighs, Fs, Hs, qF, qH, tF, tH, nearL, farL, LF, LH, LFf, LHf);
      for (int k = 0; k < ns; k++) {
        res.insert(res.end(), qF[k].begin(), qF[k].end());
        res.insert(res.end(), qH[k].begin(), qH[k].end());
        res.push_back(tF[k]);
        res.push_back(tH[k]);
      }
      for (int k = 0; k < ns; k++) {
        std::vector<double> a = padded(*LF[k], N_POLES), b = padded(*LH[k], N_POLES);
        std::vector<double> c2 = padded(*LFf[k], N_POLES), d = padded(*LHf[k], N_POLES);
        res.insert(res.end(), a.begin(), a.end());
        res.insert(res.end(), b.begin(), b.end());
        res.insert(res.end(), c2.begin(), c2.end());
        res.insert(res.end(), d.begin(), d.end());
      }
    } else if (c == "pqr") {  // path -> ncen (x y z)*ncen rad*ncen natom (x y z)*natom q*natom
      char path[4096];
      if (scanf("%4095s", path) != 1) return 3;
      std::vector<CPnt> pnt, cen;
      std::vector<REAL> ch, rad;
      readpqr(path, pnt, ch, rad, cen);
      res.push_back(cen.size());
      for (size_t i = 0; i < cen.size(); i++) {
        res.push_back(cen[i].x());
        res.push_back(cen[i].y());
        res.push_back(cen[i].z());
      }
      res.insert(res.end(), rad.begin(), rad.end());
      res.push_back(pnt.size());
      for (size_t i = 0; i < pnt.size(); i++) {
        res.push_back(pnt[i].x());
        res.push_back(pnt[i].y());
        res.push_back(pnt[i].z());
      }
      res.insert(res.end(), ch.begin(), ch.end());
    } else if (c == "wrmat") {  // path p values(p^4) -> 0
      char path[4096];
      if (scanf("%4095s", path) != 1) return 3;
      int p = ri();
      std::vector<double> m = rdv(p * p * p * p);
      writeMat_Unformatted(&m[0], p, path);
      res.push_back(0);
    } else if (c == "rdmat") {  // path maxp -> maxp^4 values
      char path[4096];
      if (scanf("%4095s", path) != 1) return 3;
      int p = ri();
      std::vector<double> m(p * p * p * p, 0.0);
      readMat_Unformatted(&m[0], p, path);
      res = m;
    } else if (c == "wrexp") {  // path p scale rcut values(p^2) -> 0
      char path[4096];
      if (scanf("%4095s", path) != 1) return 3;
      int p = ri();
      double scale = rd(), rcut = rd();
      std::vector<double> v = rdv(p * p);
      CMulExpan E(&v[0], CRange(0, p), scale);
      writeExpansion(E, rcut, path);
      res.push_back(0);
    } else if (c == "rdexp") {  // path maxp -> scale rcut values(maxp^2)
      char path[4096];
      if (scanf("%4095s", path) != 1) return 3;
      int p = ri();
      std::vector<double> v;
      REAL scale = 0, rcut = 0;
      readExpansion(v, p, path, scale, rcut);
      res.push_back(scale);
      res.push_back(rcut);
      res.insert(res.end(), v.begin(), v.end());
    } else if (c == "rdcfg") {  // path -> n (x y z w qx qy qz type)*n
      char path[4096];
      if (scanf("%4095s", path) != 1) return 3;
      std::vector<CPnt> rc;
      std::vector<CQuat> rots;
      std::vector<int> ids, types, nps;
      readConfigWithMolType(path, rc, rots, ids, types, nps);
      res.push_back(rc.size());
      for (size_t i = 0; i < rc.size(); i++) {
        res.push_back(rc[i].x());
        res.push_back(rc[i].y());
        res.push_back(rc[i].z());
        res.push_back(rots[i].w());
        res.push_back(rots[i].x());
        res.push_back(rots[i].y());
        res.push_back(rots[i].z());
        res.push_back(ids[i]);
      }
    } else if (c == "potspace") {
      // npts (x y z)*npts | molecule 0 | molecule 1 (with F, H) | pos0 q0 pos1 q1
      //   -> computePotInSpace at each point, then computeTotalIntEnergy
      int npts = ri();
      std::vector<double> P = rdv(3 * npts);
      CMolecule::N_MOL = 0;
      std::vector<CMolecule *> mols;
      mols.push_back(bd_molecule(0, true));
      mols.push_back(bd_molecule(1, true));
      for (int i = 0; i < 2; i++) {
        double x = rd(), y = rd(), z = rd();
        double w = rd(), qx = rd(), qy = rd(), qz = rd();
        mols[i]->setPos(CPnt(x, y, z));
        mols[i]->setOrient(CQuat(w, CPnt(qx, qy, qz)));
      }
      for (int t = 0; t < npts; t++)
        res.push_back(CMolecule::computePotInSpace(mols, CPnt(P[3 * t], P[3 * t + 1], P[3 * t + 2])));
      res.push_back(CMolecule::computeTotalIntEnergy(mols));
    } else if (c == "bd") {
      // seed place nsteps Dtr Dr1 Dr2 ft(9) | molecule 0 | molecule 1 | [pos0 q0 pos1 q1 if !place]
      // | ndef (ncontact npair (k0 k1 dist)*npair)*ndef
      //   -> per step: status dt pos0(3) q0(4) pos1(3) q1(4) rot1(4) rot2(4)
      // Must be the only "bd" command of its process: normRand keeps a static spare value.
      long seed = (long)rd();
      int place = ri(), nsteps = ri();
      double Dtr = rd(), Dr1 = rd(), Dr2 = rd();
      std::vector<double> ft = rdv(9);
      CMolecule::N_MOL = 0;
      std::vector<CMolecule *> mols;
      mols.push_back(bd_molecule(0));
      mols.push_back(bd_molecule(1));
      // a CBDnam around the two molecules, without the file-reading constructor
      CBDnam *bd = static_cast<CBDnam *>(::operator new(sizeof(CBDnam)));
      memset((void *)bd, 0, sizeof(CBDnam));
      new (&bd->m_mols) std::vector<CMolecule *>(mols);
      bd->m_Dtr = Dtr;
      bd->m_Dr1 = Dr1;
      bd->m_Dr2 = Dr2;
      bd->m_rot1.identity();
      bd->m_rot2.identity();
      if (place) {  // CBDnam::run, BDnam.cpp:171-181
        srand48(seed);
        CQuat Q;
        mols[0]->setPos(CPnt());
        Q = CQuat::chooseRandom();
        mols[0]->setOrient(Q);
        mols[1]->setPos(b_DIST * randOrient());
        Q = CQuat::chooseRandom();
        mols[1]->setOrient(Q);
      } else {
        for (int i = 0; i < 2; i++) {
          double x = rd(), y = rd(), z = rd();
          double w = rd(), qx = rd(), qy = rd(), qz = rd();
          mols[i]->setPos(CPnt(x, y, z));
          mols[i]->setOrient(CQuat(w, CPnt(qx, qy, qz)));
        }
        srand48(seed);
      }
      CBDnam::initConstants(2);
      CBDnam::MOLCONTACTLIST.assign(1, std::vector<CMolContact>());
      int ndef = ri();
      for (int d = 0; d < ndef; d++) {
        int ncontact = ri(), npair = ri();
        std::vector<CContact> cl;
        for (int k = 0; k < npair; k++) {
          int k0 = ri(), k1 = ri();
          double dist = rd();
          cl.push_back(CContact(k0, k1, dist));
        }
        CBDnam::MOLCONTACTLIST[0].push_back(CMolContact(1, ncontact, cl));
      }
      CPnt F1(ft[0], ft[1], ft[2]), T0(ft[3], ft[4], ft[5]), T1(ft[6], ft[7], ft[8]);
      CBDnam::STATUS status = CBDnam::RUNNING;
      REAL fact = FACT2;
      for (int st = 0; st < nsteps && status == CBDnam::RUNNING; st++) {  // CBDnam::run's loop
        REAL dt = bd->compute_dt();
        REAL dist = CMolecule::getC2CDist(mols[0], mols[1]);
        CPnt dR2, dO1, dO2;
        if (dist < f_DIST) {  // the given force and torques stand in for computeForces
          dR2 = (bd->m_Dtr * dt * fact) * F1;
          dO1 = (bd->m_Dr1 * dt * IKbT * COUL_K) * T0;
          dO2 = (bd->m_Dr2 * dt * IKbT * COUL_K) * T1;
        } else {
          dR2.zero();
          dO1.zero();
          dO2.zero();
        }
        status = bd->makeMove(dR2, dO1, dO2, dt);
        res.push_back((double)status);
        res.push_back(dt);
        for (int i = 0; i < 2; i++) {
          CPnt r = mols[i]->getRCen();
          CQuat q = mols[i]->getOrient();
          res.push_back(r.x());
          res.push_back(r.y());
          res.push_back(r.z());
          res.push_back(q.w());
          res.push_back(q.x());
          res.push_back(q.y());
          res.push_back(q.z());
        }
        const CQuat *rq[2] = {&bd->m_rot1, &bd->m_rot2};
        for (int i = 0; i < 2; i++) {
          res.push_back(rq[i]->w());
          res.push_back(rq[i]->x());
          res.push_back(rq[i]->y());
          res.push_back(rq[i]->z());
        }
      }
    } else {
      std::cout.rdbuf(old);
      fprintf(stderr, "unknown command %s\n", cmd);
      return 4;
    }
    std::cout.rdbuf(old);
    out(res);
  }
  return 0;
}
