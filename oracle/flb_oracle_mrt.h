/*
 * MRT collision of the oracle (TEST INFRASTRUCTURE, see flb_oracle.h).
 * Included by flb_oracle.c with REAL / SFX / SQRT defined, before
 * flb_oracle_impl.h.
 */

#define MCAT_(a, b) a##_##b
#define MCAT(a, b) MCAT_(a, b)
#define MFN(name) MCAT(name, SFX)

/* Forward declarations of helpers defined in flb_oracle_impl.h. */
static void MFN(rhou_cell)(int q, const REAL* c, REAL* rho, REAL* u);
static void MFN(feq_cell)(int q, const REAL* w, REAL rho, const REAL* u, REAL* feq);
static void MFN(fields)(const flbo_lattice* L, size_t id, size_t N, REAL* rho, REAL* u);
static REAL MFN(smag_omega)(int q, const REAL* c, REAL rho, const REAL* u, REAL omega, REAL cs,
                            REAL rho_field);

/* D2Q9: collision::MRT_Feq_RhoU<D2Q9> as unrolled in collisionMRT.ur.h:45-129
 * (closed-form equilibrium moments, rates {0, 1.1, 1.1, 0, 1.1, 0, 1.1, w, w}).
 * D3Q19: m = M f, m_eq = M feq(SecondOrder), f <- f - M^T diag(s_j / |M_j|^2)
 * (m - m_eq) on the orthogonal d'Humieres basis (no reference form; the
 * factors s_j / |M_j|^2 are rounded as the GPU kernel rounds them, and both
 * evaluate the transforms on the pairs of opposite directions). */
static void MFN(mrt_cell)(int q, const REAL* w, REAL omega, const flbo_task* t,
                          const flbo_lattice* L, REAL* c, size_t id, size_t N) {
  REAL rho, u[3] = {0, 0, 0};
  MFN(rhou_cell)(q, c, &rho, u);
  REAL om = omega;
  if (t->dynamics == FLBO_DYN_SMAG_MRT) {
    REAL rf, uf[3];
    MFN(fields)(L, id, N, &rf, uf);
    om = MFN(smag_omega)(q, c, rho, u, omega, (REAL)t->smagorinsky, rf);
  }
  if (q == 9) {
    REAL rt[9];
    if (L->mrt_rates) {
      for (int j = 0; j < 9; ++j) rt[j] = isnan(L->mrt_rates[j]) ? om : (REAL)L->mrt_rates[j];
    } else {
      const REAL s11 = (REAL)(11. / 10.);
      const REAL d[9] = {0, s11, s11, 0, s11, 0, s11, om, om};
      for (int j = 0; j < 9; ++j) rt[j] = d[j];
    }
    REAL ims[9][9];
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j)
        ims[i][j] = ((REAL)INVM9_N[i][j] / (REAL)INVM9_D[i][j]) * rt[j];
    REAL m[9];
    for (int i = 0; i < 9; ++i) {
      m[i] = (REAL)0;
      for (int j = 0; j < 9; ++j) m[i] += (REAL)M9[i][j] * c[j];
    }
    REAL dm[9];
    const REAL ux2 = u[0] * u[0];
    const REAL uy2 = u[1] * u[1];
    const REAL rhoux = rho * u[0];
    const REAL rhouy = rho * u[1];
    const REAL cse0 = (REAL)3 * rho * (ux2 + uy2);
    dm[0] = m[0] - rho;
    dm[1] = m[1] + (REAL)2 * rho - cse0;
    dm[2] = m[2] - (REAL)9 * rho * ux2 * uy2 + cse0 - rho;
    dm[3] = m[3] - rhoux;
    dm[4] = m[4] - rhoux * ((REAL)3 * uy2 - (REAL)1);
    dm[5] = m[5] - rhouy;
    dm[6] = m[6] - rhouy * ((REAL)3 * ux2 - (REAL)1);
    dm[7] = m[7] - rho * (ux2 - uy2);
    dm[8] = m[8] - rho * u[0] * u[1];
    for (int i = 0; i < 9; ++i) {
      REAL s = (REAL)0;
      for (int j = 0; j < 9; ++j) s += ims[i][j] * dm[j];
      c[i] -= s;
    }
  } else {
    /* On the pairs of opposite directions (i = 2p+1, i+1 = opp i): with
       n = f - feq, s_p = n_i + n_{i+1}, d_p = n_i - n_{i+1}, an even row is
       M(j,0) n_0 + sum_p M(j,i) s_p and an odd row sum_p M(j,i) d_p; back,
       E_p / O_p sum M(j,i) k_j over the even / odd rows, f_i -= E_p + O_p,
       f_{i+1} -= E_p - O_p. Each sum starts from zero and runs over the
       nonzero entries in ascending index. */
    REAL feq[19], sp[9], dp[9], k[19];
    MFN(feq_cell)(q, w, rho, u, feq);
    const REAL n0 = c[0] - feq[0];
    for (int p = 0; p < 9; ++p) {
      const REAL a = c[2 * p + 1] - feq[2 * p + 1];
      const REAL b = c[2 * p + 2] - feq[2 * p + 2];
      sp[p] = a + b;
      dp[p] = a - b;
    }
    for (int j = 0; j < 19; ++j) {
      /* the row's factor s_j / |M_j|^2: a fixed rate over the norm in double,
         rounded once; a shear row (NaN or negative rate) the cell's omega times
         the rounded reciprocal of the norm */
      double sj = L->mrt_rates ? L->mrt_rates[j] : mrt19_default_rate(j);
      const REAL sn = (isnan(sj) || sj < 0.0) ? om * (REAL)(1.0 / (double)M19_NORM[j])
                                              : (REAL)(sj / (double)M19_NORM[j]);
      REAL dm = (REAL)0;
      if (!M19_ODD[j] && M19[j][0] != 0) dm += (REAL)M19[j][0] * n0;
      for (int p = 0; p < 9; ++p) {
        const int m = M19[j][2 * p + 1];
        if (m != 0) dm += (REAL)m * (M19_ODD[j] ? dp[p] : sp[p]);
      }
      k[j] = dm * sn;
    }
    REAL d0 = (REAL)0;
    for (int j = 0; j < 19; ++j)
      if (!M19_ODD[j] && M19[j][0] != 0) d0 += (REAL)M19[j][0] * k[j];
    c[0] -= d0;
    for (int p = 0; p < 9; ++p) {
      REAL E = (REAL)0, O = (REAL)0;
      for (int j = 0; j < 19; ++j) {
        const int m = M19[j][2 * p + 1];
        if (m == 0) continue;
        if (M19_ODD[j]) O += (REAL)m * k[j];
        else E += (REAL)m * k[j];
      }
      c[2 * p + 1] -= E + O;
      c[2 * p + 2] -= E - O;
    }
  }
}

#undef MFN
#undef MCAT
#undef MCAT_
