/*
 * Precision-generic body of the oracle (TEST INFRASTRUCTURE, see flb_oracle.h).
 * Included twice by flb_oracle.c with REAL = float / double and SFX = f32 / f64.
 *
 * Expression order follows the reference's unrolled specialisations
 * (compiled with -D_UNROLLFOR, as benchmarks/cavity3d/Makefile:9 does), so
 * that sums and products round exactly as FreeLB's do.
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* Lattice weights as the reference computes them: T(1)/T(3) etc.
 * (lattice_set.h:168-194, 259-296). */
static void FN(weights)(int q, REAL* w) {
  if (q == 19) {
    w[0] = (REAL)1 / (REAL)3;
    for (int i = 1; i <= 6; ++i) w[i] = (REAL)1 / (REAL)18;
    for (int i = 7; i <= 18; ++i) w[i] = (REAL)1 / (REAL)36;
  } else if (q == 9) {
    w[0] = (REAL)4 / (REAL)9;
    for (int i = 1; i <= 4; ++i) w[i] = (REAL)1 / (REAL)9;
    for (int i = 5; i <= 8; ++i) w[i] = (REAL)1 / (REAL)36;
  } else if (q == 5) {
    /* D2Q5: T(1)/T(3), T(1)/T(6) (lattice_set.h:146-150) */
    w[0] = (REAL)1 / (REAL)3;
    for (int i = 1; i <= 4; ++i) w[i] = (REAL)1 / (REAL)6;
  } else {
    /* D3Q7: T(1)/T(4), T(1)/T(8) (lattice_set.h:198-202) */
    w[0] = (REAL)1 / (REAL)4;
    for (int i = 1; i <= 6; ++i) w[i] = (REAL)1 / (REAL)8;
  }
}

/* moment::rhoImpl<D2Q5 / D3Q7> (moment.ur.h:47-82): left-to-right sum. */
static REAL FN(scalar_sum)(int q, const REAL* c) {
  REAL r = c[0];
  for (int i = 1; i < q; ++i) r = r + c[i];
  return r;
}

/* moment::rhoUImpl<D3Q19> (moment.ur.h:505-517) and <D2Q9> (moment.ur.h:461-472). */
static void FN(rhou_cell)(int q, const REAL* c, REAL* rho, REAL* u) {
  if (q == 19) {
    REAL r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8] + c[9] + c[10] + c[11] +
             c[12] + c[13] + c[14] + c[15] + c[16] + c[17] + c[18];
    *rho = r;
    u[0] = (c[1] - c[2] + c[7] - c[8] + c[9] - c[10] + c[13] - c[14] + c[15] - c[16]) / r;
    u[1] = (c[3] - c[4] + c[7] - c[8] + c[11] - c[12] - c[13] + c[14] + c[17] - c[18]) / r;
    u[2] = (c[5] - c[6] + c[9] - c[10] + c[11] - c[12] - c[15] + c[16] - c[17] + c[18]) / r;
  } else if (q == 9) {
    REAL r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8];
    *rho = r;
    u[0] = (c[1] - c[2] + c[5] - c[6] + c[7] - c[8]) / r;
    u[1] = (c[3] - c[4] + c[5] - c[6] - c[7] + c[8]) / r;
  } else {
    /* scalar lattices: the sum and (sum c f) / sum as a read-out */
    const REAL r = FN(scalar_sum)(q, c);
    *rho = r;
    for (int a = 0; a < lat_d(q); ++a) u[a] = (c[2 * a + 1] - c[2 * a + 2]) / r;
  }
}

/* moment::forcerhoU (forcerhoUImpl<D3Q19> moment.ur.h:639-663, <D2Q9> 564-574):
 * u = (sum c f + F/2) / rho, each component divided separately. */
static void FN(force_rhou_cell)(int q, const REAL* c, const REAL* F, REAL* rho, REAL* u) {
  if (q == 19) {
    REAL r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8] + c[9] + c[10] + c[11] +
             c[12] + c[13] + c[14] + c[15] + c[16] + c[17] + c[18];
    *rho = r;
    u[0] = (c[1] - c[2] + c[7] - c[8] + c[9] - c[10] + c[13] - c[14] + c[15] - c[16] + F[0] * (REAL)0.5) / r;
    u[1] = (c[3] - c[4] + c[7] - c[8] + c[11] - c[12] - c[13] + c[14] + c[17] - c[18] + F[1] * (REAL)0.5) / r;
    u[2] = (c[5] - c[6] + c[9] - c[10] + c[11] - c[12] - c[15] + c[16] - c[17] + c[18] + F[2] * (REAL)0.5) / r;
  } else {
    REAL r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8];
    *rho = r;
    u[0] = (c[1] - c[2] + c[5] - c[6] + c[7] - c[8] + F[0] * (REAL)0.5) / r;
    u[1] = (c[3] - c[4] + c[5] - c[6] - c[7] + c[8] + F[1] * (REAL)0.5) / r;
  }
}

/* force::ForcePopImpl<D3Q19> / <D2Q9> (force.ur.h, the unrolled form of
 * force.h:29-36): Fi = w_i InvCs2 (F.(c_i - u) + InvCs2 (c_i.u)(c_i.F)),
 * written per direction with the terms c_ia - u_a + InvCs2 (c_i.u) c_ia. */
static void FN(force_pop_cell)(int q, const REAL* w, const REAL* u, const REAL* F, REAL* Fi) {
  const REAL I2 = (REAL)3;
  const REAL u0 = u[0], u1 = u[1], F0 = F[0], F1 = F[1];
  const REAL T1_u0 = (REAL)1 - u0, _T1_u0 = (REAL)-1 - u0, _u0 = -u0;
  const REAL T1_u1 = (REAL)1 - u1, _T1_u1 = (REAL)-1 - u1, _u1 = -u1;
  if (q == 19) {
    const REAL u2 = u[2], F2 = F[2];
    const REAL T1_u2 = (REAL)1 - u2, _T1_u2 = (REAL)-1 - u2, _u2 = -u2;
    const REAL c1 = I2 * (u0), c3 = I2 * (u1), c5 = I2 * (u2), c7 = I2 * (u0 + u1),
               c9 = I2 * (u0 + u2), c11 = I2 * (u1 + u2), c13 = I2 * (u0 - u1),
               c15 = I2 * (u0 - u2), c17 = I2 * (u1 - u2);
    Fi[0] = w[0] * I2 * (-(F0 * u0 + F1 * u1 + F2 * u2));
    Fi[1] = w[1] * I2 * (F0 * (T1_u0 + c1) + F1 * (_u1) + F2 * (_u2));
    Fi[2] = w[2] * I2 * (F0 * (_T1_u0 + c1) + F1 * (_u1) + F2 * (_u2));
    Fi[3] = w[3] * I2 * (F0 * (_u0) + F1 * (T1_u1 + c3) + F2 * (_u2));
    Fi[4] = w[4] * I2 * (F0 * (_u0) + F1 * (_T1_u1 + c3) + F2 * (_u2));
    Fi[5] = w[5] * I2 * (F0 * (_u0) + F1 * (_u1) + F2 * (T1_u2 + c5));
    Fi[6] = w[6] * I2 * (F0 * (_u0) + F1 * (_u1) + F2 * (_T1_u2 + c5));
    Fi[7] = w[7] * I2 * (F0 * (T1_u0 + c7) + F1 * (T1_u1 + c7) + F2 * (_u2));
    Fi[8] = w[8] * I2 * (F0 * (_T1_u0 + c7) + F1 * (_T1_u1 + c7) + F2 * (_u2));
    Fi[9] = w[9] * I2 * (F0 * (T1_u0 + c9) + F1 * (_u1) + F2 * (T1_u2 + c9));
    Fi[10] = w[10] * I2 * (F0 * (_T1_u0 + c9) + F1 * (_u1) + F2 * (_T1_u2 + c9));
    Fi[11] = w[11] * I2 * (F0 * (_u0) + F1 * (T1_u1 + c11) + F2 * (T1_u2 + c11));
    Fi[12] = w[12] * I2 * (F0 * (_u0) + F1 * (_T1_u1 + c11) + F2 * (_T1_u2 + c11));
    Fi[13] = w[13] * I2 * (F0 * (T1_u0 + c13) + F1 * (_T1_u1 - c13) + F2 * (_u2));
    Fi[14] = w[14] * I2 * (F0 * (_T1_u0 + c13) + F1 * (T1_u1 - c13) + F2 * (_u2));
    Fi[15] = w[15] * I2 * (F0 * (T1_u0 + c15) + F1 * (_u1) + F2 * (_T1_u2 - c15));
    Fi[16] = w[16] * I2 * (F0 * (_T1_u0 + c15) + F1 * (_u1) + F2 * (T1_u2 - c15));
    Fi[17] = w[17] * I2 * (F0 * (_u0) + F1 * (T1_u1 + c17) + F2 * (_T1_u2 - c17));
    Fi[18] = w[18] * I2 * (F0 * (_u0) + F1 * (_T1_u1 + c17) + F2 * (T1_u2 - c17));
  } else {
    const REAL c1 = I2 * (u0), c3 = I2 * (u1), c5 = I2 * (u0 + u1), c7 = I2 * (u0 - u1);
    Fi[0] = w[0] * I2 * (-(F0 * u0 + F1 * u1));
    Fi[1] = w[1] * I2 * (F0 * (T1_u0 + c1) + F1 * (_u1));
    Fi[2] = w[2] * I2 * (F0 * (_T1_u0 + c1) + F1 * (_u1));
    Fi[3] = w[3] * I2 * (F0 * (_u0) + F1 * (T1_u1 + c3));
    Fi[4] = w[4] * I2 * (F0 * (_u0) + F1 * (_T1_u1 + c3));
    Fi[5] = w[5] * I2 * (F0 * (T1_u0 + c5) + F1 * (T1_u1 + c5));
    Fi[6] = w[6] * I2 * (F0 * (_T1_u0 + c5) + F1 * (_T1_u1 + c5));
    Fi[7] = w[7] * I2 * (F0 * (T1_u0 + c7) + F1 * (_T1_u1 - c7));
    Fi[8] = w[8] * I2 * (F0 * (_T1_u0 + c7) + F1 * (T1_u1 - c7));
  }
}

/* forcerhoUImpl's scalar-force overload (moment.h:423-435; unrolled
 * moment.ur.h:574-583 D2Q9, 652-663 D3Q19): sums, F/2 on the last axis, then
 * every component divided by rho. */
static void FN(scalar_force_rhou_cell)(int q, const REAL* c, REAL F, REAL* rho, REAL* u) {
  if (q == 19) {
    REAL r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8] + c[9] + c[10] + c[11] +
             c[12] + c[13] + c[14] + c[15] + c[16] + c[17] + c[18];
    *rho = r;
    u[0] = c[1] - c[2] + c[7] - c[8] + c[9] - c[10] + c[13] - c[14] + c[15] - c[16];
    u[1] = c[3] - c[4] + c[7] - c[8] + c[11] - c[12] - c[13] + c[14] + c[17] - c[18];
    u[2] = c[5] - c[6] + c[9] - c[10] + c[11] - c[12] - c[15] + c[16] - c[17] + c[18];
    u[2] += F * (REAL)0.5;
    u[0] /= r;
    u[1] /= r;
    u[2] /= r;
  } else {
    REAL r = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8];
    *rho = r;
    u[0] = c[1] - c[2] + c[5] - c[6] + c[7] - c[8];
    u[1] = c[3] - c[4] + c[5] - c[6] - c[7] + c[8];
    u[1] += F * (REAL)0.5;
    u[0] /= r;
    u[1] /= r;
  }
}

/* ScalarForcePopImpl<D2Q9, 1> (force.ur.h:332-353) and <D3Q19, 2>
 * (force.ur.h:584-616): Fi = w_i InvCs2 F (c_id - u_d + InvCs2 (c_i.u) c_id). */
static void FN(scalar_force_pop_cell)(int q, const REAL* w, const REAL* u, REAL F, REAL* Fi) {
  const REAL I2 = (REAL)3;
  if (q == 19) {
    const REAL u0 = u[0], u1 = u[1], u2 = u[2];
    const REAL T1 = (REAL)1 - u2, _T1 = (REAL)-1 - u2, _u = -u2;
    const REAL k5 = I2 * (u2), k9 = I2 * (u0 + u2), k11 = I2 * (u1 + u2), k15 = I2 * (u0 - u2),
               k17 = I2 * (u1 - u2);
    const REAL t[19] = {_u, _u, _u, _u, _u, T1 + k5, _T1 + k5, _u, _u, T1 + k9, _T1 + k9, T1 + k11, _T1 + k11,
                        _u, _u, _T1 - k15, T1 - k15, _T1 - k17, T1 - k17};
    for (int i = 0; i < 19; ++i) Fi[i] = w[i] * I2 * F * (t[i]);
  } else {
    const REAL u0 = u[0], u1 = u[1];
    const REAL T1 = (REAL)1 - u1, _T1 = (REAL)-1 - u1, _u = -u1;
    const REAL k3 = I2 * (u1), k5 = I2 * (u0 + u1), k7 = I2 * (u0 - u1);
    const REAL t[9] = {_u, _u, _u, T1 + k3, _T1 + k3, T1 + k5, _T1 + k5, _T1 - k7, T1 - k7};
    for (int i = 0; i < 9; ++i) Fi[i] = w[i] * I2 * F * (t[i]);
  }
}

/* equilibrium::SecondOrderImpl<D3Q19> (equilibrium.ur.h:165-214) and <D2Q9>
 * (equilibrium.ur.h:65-94). */
static void FN(feq_cell)(int q, const REAL* w, REAL rho, const REAL* u, REAL* feq) {
  const REAL InvCs2 = (REAL)3;
  const REAL InvCs2x = (REAL)0.5 * InvCs2;
  if (q == 5 || q == 7) {
    /* SecondOrderImpl<D2Q5> (equilibrium.ur.h:43-63), <D3Q7> (96-119) */
    const REAL var0 = q == 7 ? (REAL)1 - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * InvCs2x
                             : (REAL)1 - (u[0] * u[0] + u[1] * u[1]) * InvCs2x;
    const REAL rhowk0 = w[0] * rho;
    const REAL rhowk1 = w[1] * rho;
    feq[0] = rhowk0 * var0;
    for (int a = 0; a < lat_d(q); ++a) {
      const REAL uc = InvCs2 * (u[a]);
      const REAL v = var0 + uc * uc * (REAL)0.5;
      feq[2 * a + 1] = rhowk1 * (v + uc);
      feq[2 * a + 2] = rhowk1 * (v - uc);
    }
    return;
  }
  if (q == 19) {
    const REAL var0 = (REAL)1 - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * InvCs2x;
    const REAL rhowk0 = w[0] * rho;
    const REAL rhowk1 = w[1] * rho;
    const REAL rhowk7 = w[7] * rho;
    /* c.u for the first direction of each opposite pair 1,3,5,7,9,11,13,15,17 */
    const REAL cu[9] = {u[0],        u[1],        u[2],        u[0] + u[1], u[0] + u[2],
                        u[1] + u[2], u[0] - u[1], u[0] - u[2], u[1] - u[2]};
    feq[0] = rhowk0 * var0;
    for (int p = 0; p < 9; ++p) {
      const REAL a = InvCs2 * cu[p];
      const REAL v = var0 + a * a * (REAL)0.5;
      const REAL rw = p < 3 ? rhowk1 : rhowk7;
      feq[2 * p + 1] = rw * (v + a);
      feq[2 * p + 2] = rw * (v - a);
    }
  } else {
    const REAL var0 = (REAL)1 - (u[0] * u[0] + u[1] * u[1]) * InvCs2x;
    const REAL rhowk0 = w[0] * rho;
    const REAL rhowk1 = w[1] * rho;
    const REAL rhowk5 = w[5] * rho;
    const REAL cu[4] = {u[0], u[1], u[0] + u[1], u[0] - u[1]};
    feq[0] = rhowk0 * var0;
    for (int p = 0; p < 4; ++p) {
      const REAL a = InvCs2 * cu[p];
      const REAL v = var0 + a * a * (REAL)0.5;
      const REAL rw = p < 2 ? rhowk1 : rhowk5;
      feq[2 * p + 1] = rw * (v + a);
      feq[2 * p + 2] = rw * (v - a);
    }
  }
}

/* c_i . u for direction i, summed over axes in order from zero the way
 * Vector<T,D>::operator* does (Vector.h:229-236). */
static REAL FN(cdotu)(int q, int i, const REAL* u) {
  const int* ci = &lat_c(q)[i][0];
  const int d = lat_d(q);
  REAL r = (REAL)0;
  for (int a = 0; a < d; ++a) r += u[a] * (REAL)ci[a];
  return r;
}

/* Fetch RHO / VELOCITY field values of a cell (fields default to rho = 1, u = 0). */
static void FN(fields)(const flbo_lattice* L, size_t id, size_t N, REAL* rho, REAL* u) {
  const int d = lat_d(L->q);
  *rho = L->rho_field ? (REAL)L->rho_field[id] : (REAL)1;
  for (int a = 0; a < d; ++a) u[a] = L->u_field ? (REAL)L->u_field[(size_t)a * N + id] : (REAL)0;
}

/* Smagorinsky effective omega (collisionLES.h:36-59): Pi_ab_neq from
 * moment.ur.h:708-718 (D2Q9) / 750-763 (D3Q19), NormSquare from util.h:418-434. */
static REAL FN(smag_omega)(int q, const REAL* c, REAL rho, const REAL* u, REAL omega, REAL cs,
                           REAL rho_field) {
  const REAL cs2 = (REAL)1 / (REAL)3;
  REAL ns;
  if (q == 19) {
    REAL t0 = c[1] + c[2] + c[7] + c[8] + c[9] + c[10] + c[13] + c[14] + c[15] + c[16] -
              rho * (u[0] * u[0] + cs2);
    REAL t1 = c[7] + c[8] - c[13] - c[14] - rho * (u[0] * u[1]);
    REAL t2 = c[9] + c[10] - c[15] - c[16] - rho * (u[0] * u[2]);
    REAL t3 = c[3] + c[4] + c[7] + c[8] + c[11] + c[12] + c[13] + c[14] + c[17] + c[18] -
              rho * (u[1] * u[1] + cs2);
    REAL t4 = c[11] + c[12] - c[17] - c[18] - rho * (u[1] * u[2]);
    REAL t5 = c[5] + c[6] + c[9] + c[10] + c[11] + c[12] + c[15] + c[16] + c[17] + c[18] -
              rho * (u[2] * u[2] + cs2);
    ns = t0 * t0 + t1 * t1 * 2 + t2 * t2 * 2 + t3 * t3 + t4 * t4 * 2 + t5 * t5;
  } else {
    REAL t0 = c[1] + c[2] + c[5] + c[6] + c[7] + c[8] - rho * (u[0] * u[0] + cs2);
    REAL t1 = c[5] + c[6] - c[7] - c[8] - rho * (u[0] * u[1]);
    REAL t2 = c[3] + c[4] + c[5] + c[6] + c[7] + c[8] - rho * (u[1] * u[1] + cs2);
    ns = t0 * t0 + t1 * t1 * 2 + t2 * t2;
  }
  const REAL norm = SQRT(ns);
  /* smagorinsky * smagorinsky * InvCs2 * InvCs2 * 2 * std::sqrt(2) / RHO_field;
   * std::sqrt(2) of an int argument is a double, so the product is formed in
   * double from that point on and rounded back to REAL on assignment. */
  const REAL coeff = (REAL)((double)(cs * cs * (REAL)3 * (REAL)3 * (REAL)2) * sqrt(2.0) /
                            (double)rho_field);
  const REAL rt_mol = (REAL)1 / omega;
  const REAL rt_tur = (REAL)0.5 * (SQRT(rt_mol * rt_mol + coeff * norm) - rt_mol);
  return (REAL)1 / (rt_mol + rt_tur);
}

/* Apply one cell's dynamics in place on the gathered values c[q]. */
static void FN(collide_cell)(const flbo_lattice* L, const flbo_task* t, const REAL* w, REAL omega,
                             REAL _omega, REAL* c, size_t id, size_t N) {
  const int q = L->q;
  const int d = lat_d(q);
  REAL rho, u[3] = {0, 0, 0}, feq[19];
  switch (t->dynamics) {
    case FLBO_DYN_BGK:
    case FLBO_DYN_BGK_FIELD:
    case FLBO_DYN_BGK_WRITE:
      /* collision::BGK (collision.h:39-56) */
      if (t->dynamics == FLBO_DYN_BGK_FIELD) {
        FN(fields)(L, id, N, &rho, u);
      } else {
        FN(rhou_cell)(q, c, &rho, u);
        if (t->dynamics == FLBO_DYN_BGK_WRITE) {
          /* rhoUImpl<CELL, true> (moment.h:373-376) */
          if (L->rho_field) L->rho_field[id] = (double)rho;
          if (L->u_field)
            for (int a = 0; a < d; ++a) L->u_field[(size_t)a * N + id] = (double)u[a];
        }
      }
      FN(feq_cell)(q, w, rho, u, feq);
      for (int i = 0; i < q; ++i) c[i] = omega * feq[i] + _omega * c[i];
      break;
    case FLBO_DYN_BOUNCEBACK:
      /* collision::BounceBack<D3Q19/D2Q9> (collision.ur.h:44-93) */
      for (int i = 1; i < q; i += 2) {
        const REAL tmp = c[i];
        c[i] = c[i + 1];
        c[i + 1] = tmp;
      }
      break;
    case FLBO_DYN_BOUNCEBACK_MW: {
      /* collision::BounceBackMovingWall (collision.ur.h:97-187):
       * rhox = 2*InvCs2*RHO; uc_i = (c_i.u) * w_i * rhox */
      REAL rf, uf[3] = {0, 0, 0};
      FN(fields)(L, id, N, &rf, uf);
      const REAL rhox = (REAL)6 * rf;
      for (int i = 1; i < q; i += 2) {
        const int* ci = &lat_c(q)[i][0];
        /* the unrolled form sums only the non-zero components: u0, u1, u0+u1, u0-u1, ... */
        REAL cu = 0;
        int first = 1;
        for (int a = 0; a < d; ++a) {
          if (ci[a] == 0) continue;
          if (first) {
            cu = ci[a] > 0 ? uf[a] : -uf[a];
            first = 0;
          } else {
            cu = ci[a] > 0 ? cu + uf[a] : cu - uf[a];
          }
        }
        const REAL uc = cu * w[i] * rhox;
        const REAL tmp = c[i];
        c[i] = c[i + 1] + uc;
        c[i + 1] = tmp - uc;
      }
      break;
    }
    case FLBO_DYN_SMAG_BGK: {
      /* collision::SmagorinskyBGK (collisionLES.h:36-66); RHO field in the coefficient */
      REAL rf, uf[3];
      FN(fields)(L, id, N, &rf, uf);
      FN(rhou_cell)(q, c, &rho, u);
      const REAL eo = FN(smag_omega)(q, c, rho, u, omega, (REAL)t->smagorinsky, rf);
      FN(feq_cell)(q, w, rho, u, feq);
      const REAL _eo = (REAL)1 - eo;
      for (int i = 0; i < q; ++i) c[i] = eo * feq[i] + _eo * c[i];
      break;
    }
    case FLBO_DYN_MRT:
    case FLBO_DYN_SMAG_MRT:
      FN(mrt_cell)(q, w, omega, t, L, c, id, N);
      break;
    case FLBO_DYN_BGK_FORCE: {
      /* collision::BGKForce (collision.h:68-87): forcerhoU, ForcePop, feq, then
       * omega feq + (1 - omega) f + (1 - omega/2) Fi (fOmega, block_lattice.hh:33-34) */
      const REAL F[3] = {(REAL)L->force[0], (REAL)L->force[1], (REAL)L->force[2]};
      REAL Fi[19];
      FN(force_rhou_cell)(q, c, F, &rho, u);
      FN(force_pop_cell)(q, w, u, F, Fi);
      FN(feq_cell)(q, w, rho, u, feq);
      const REAL fomega = (REAL)1 - (REAL)0.5 * omega;
      for (int i = 0; i < q; ++i) c[i] = omega * feq[i] + _omega * c[i] + fomega * Fi[i];
      break;
    }
    case FLBO_DYN_BGK_SCALAR_FORCE: {
      /* collision::BGKForce<forcerhoU<CELL, ScalarForce, true>, SecondOrder,
       * ScalarForce> (collision.h:68-87): F is the cell's SCALARFORCE field;
       * rho, u go to the RHO / VELOCITY fields (moment.h:430-434) */
      const REAL F = L->sforce ? (REAL)L->sforce[id] : (REAL)0;
      REAL Fi[19];
      FN(scalar_force_rhou_cell)(q, c, F, &rho, u);
      if (L->rho_field) L->rho_field[id] = (double)rho;
      if (L->u_field)
        for (int a = 0; a < d; ++a) L->u_field[(size_t)a * N + id] = (double)u[a];
      FN(scalar_force_pop_cell)(q, w, u, F, Fi);
      FN(feq_cell)(q, w, rho, u, feq);
      const REAL fomega = (REAL)1 - (REAL)0.5 * omega;
      for (int i = 0; i < q; ++i) c[i] = omega * feq[i] + _omega * c[i] + fomega * Fi[i];
      break;
    }
    case FLBO_DYN_SCALAR_BGK:
    case FLBO_DYN_SCALAR_BGK_FIELD: {
      /* collision::BGK on a D2Q5 / D3Q7 lattice with u from the VELOCITY field
       * (useFieldU, moment.h:56-69); the scalar is the population sum written
       * to the field (rho<CELL, true>, moment.h:159-182) or the field value
       * (useFieldrhoU, moment.h:71-90) */
      const double* ue = L->u_ext ? L->u_ext : L->u_field;
      for (int a = 0; a < d; ++a) u[a] = ue ? (REAL)ue[(size_t)a * N + id] : (REAL)0;
      if (t->dynamics == FLBO_DYN_SCALAR_BGK) {
        rho = FN(scalar_sum)(q, c);
        if (L->rho_field) L->rho_field[id] = (double)rho;
      } else {
        rho = L->rho_field ? (REAL)L->rho_field[id] : (REAL)1;
      }
      FN(feq_cell)(q, w, rho, u, feq);
      for (int i = 0; i < q; ++i) c[i] = omega * feq[i] + _omega * c[i];
      break;
    }
    default:
      break;
  }
}

/* One reference time step on the SoA arrays A (in: post-stream state of the
 * previous step; out: post-stream state of this step). B is scratch.
 *   1. BlockLattice::ApplyCellDynamics (block_lattice.hh:72-92): every cell
 *      id in [0, N), first-match task (tmp.h:283-309), in place.
 *   2. BlockLattice::Stream (block_lattice.hh:64-70) via CyclicArray::rotate
 *      (field.h:676-685): arr_i[id] <- old arr_i[(id - Delta_i) mod N],
 *      Delta_i = c_i . (1, Nx, Nx*Ny) (block_lattice_base.h:126-127).
 *   3. BlockBoundaryManager::Apply (basic_boundary.h:345-347) for each
 *      halfway manager in order (boundary.h:121-128): for every boundary cell
 *      of the inner region and every k with a void neighbour at id + Delta_k,
 *      f_{opp k}(id) <- getPrevious_k(id) [+ 2 InvCs2 w_o RHO (VELOCITY . c_o)]
 *      (bounce_back_boundary.h:169-181, 240-256; getPrevious at field.h:645-656).
 */
static void FN(step)(const flbo_lattice* L, REAL* A, REAL* B) {
  const int q = L->q;
  const size_t nx = (size_t)L->nx, ny = (size_t)L->ny, nz = (size_t)L->nz;
  const size_t N = nx * ny * nz;
  const int (*C)[3] = lat_c(q);
  const int* OPP = lat_opp(q);
  REAL w[19];
  FN(weights)(q, w);
  const REAL omega = (REAL)L->omega;
  const REAL _omega = (REAL)1 - omega;
  const uint8_t* flags = L->flags;

#pragma omp parallel for schedule(static) num_threads(L->nthreads > 0 ? L->nthreads : omp_get_max_threads())
  for (size_t id = 0; id < N; ++id) {
    const uint8_t fl = flags[id];
    const flbo_task* t = NULL;
    for (int k = 0; k < L->ntasks; ++k) {
      if (fl & L->tasks[k].mask) {
        t = &L->tasks[k];
        break;
      }
    }
    if (!t || t->dynamics == FLBO_DYN_NONE) continue;
    REAL c[19];
    for (int i = 0; i < q; ++i) c[i] = A[(size_t)i * N + id];
    FN(collide_cell)(L, t, w, omega, _omega, c, id, N);
    for (int i = 0; i < q; ++i) A[(size_t)i * N + id] = c[i];
  }

  ptrdiff_t delta[19];
  for (int i = 0; i < q; ++i)
    delta[i] = (ptrdiff_t)C[i][0] + (ptrdiff_t)C[i][1] * (ptrdiff_t)nx +
               (ptrdiff_t)C[i][2] * (ptrdiff_t)(nx * ny);

#pragma omp parallel for schedule(static) num_threads(L->nthreads > 0 ? L->nthreads : omp_get_max_threads())
  for (size_t id = 0; id < N; ++id) {
    for (int i = 0; i < q; ++i) {
      ptrdiff_t src = (ptrdiff_t)id - delta[i];
      if (src < 0) src += (ptrdiff_t)N;
      if (src >= (ptrdiff_t)N) src -= (ptrdiff_t)N;
      B[(size_t)i * N + id] = A[(size_t)i * N + (size_t)src];
    }
  }

  const int ov = L->overlap;
  const int d = lat_d(q);
  for (int b = 0; b < L->nbc; ++b) {
    const flbo_hwbc* bc = &L->bcs[b];
    const size_t zlo = d == 3 ? (size_t)ov : 0, zhi = d == 3 ? nz - (size_t)ov : 1;
#pragma omp parallel for schedule(static) num_threads(L->nthreads > 0 ? L->nthreads : omp_get_max_threads())
    for (size_t z = zlo; z < zhi; ++z) {
      for (size_t y = (size_t)ov; y < ny - (size_t)ov; ++y) {
        for (size_t x = (size_t)ov; x < nx - (size_t)ov; ++x) {
          const size_t id = x + nx * (y + ny * z);
          if (!(flags[id] & bc->cell_mask)) continue;
          REAL rf = 1, uf[3] = {0, 0, 0};
          if (bc->kind != FLBO_HWBC_NORMAL) FN(fields)(L, id, N, &rf, uf);
          for (int k = 1; k < q; ++k) {
            const size_t nb = (size_t)((ptrdiff_t)id + delta[k]);
            if (!(flags[nb] & bc->void_mask)) continue;
            const int o = OPP[k];
            REAL v = A[(size_t)k * N + id];
            if (bc->kind == FLBO_HWBC_MOVING) {
              /* 2 * InvCs2 * w_o * RHO * (VELOCITY . c_o) */
              v = v + (REAL)6 * w[o] * rf * FN(cdotu)(q, o, uf);
            } else if (bc->kind == FLBO_HWBC_ANTI) {
              /* bounceback::anti_simple (bounce_back_boundary.h:183-193):
               * 2 * RHO * w_o - getPrevious(k) */
              v = (REAL)2 * rf * w[o] - v;
            }
            B[(size_t)o * N + id] = v;
          }
        }
      }
    }
  }
}

int FN(run_impl)(const flbo_lattice* L, REAL* pops, REAL* scratch, int steps) {
  if (!L || !pops || !scratch || !lat_ok(L->q) || steps < 0) return -1;
  const size_t N = (size_t)L->nx * (size_t)L->ny * (size_t)L->nz;
  REAL* A = pops;
  REAL* B = scratch;
  for (int s = 0; s < steps; ++s) {
    FN(step)(L, A, B);
    REAL* t = A;
    A = B;
    B = t;
  }
  if (A != pops) memcpy(pops, A, sizeof(REAL) * (size_t)L->q * N);
  return 0;
}

/* The coupled loop of examples/cazsblock2d/cazsblock2d.cpp:385-409 for one
 * flow lattice and its scalar lattices, per step:
 *   SCALARFORCE.InitValue(0) (field_struct.h:238);
 *   for each coupling: BlockLatManagerCoupling::ApplyCellDynamics with
 *     Key_TypePair<mask, force::Buoyancy> (block_lattice.h:488-509,
 *     force.h:173-181): SCALARFORCE += gbeta (RHO_scalar - rho_init) on the
 *     cells whose flag matches;
 *   the flow step (collide, stream, halfway managers);
 *   each scalar lattice's step (its u_ext is the flow's VELOCITY field).
 * pops / scratch as flbo_run, per lattice. */
int FN(run_coupled_impl)(const flbo_lattice* flow, REAL* pf, REAL* sf, const flbo_lattice* const* scal,
                         REAL* const* ps, REAL* const* ss, const flbo_buoyancy* bu, int nb, int ns, int steps) {
  if (!flow || !pf || !sf || !flow->sforce || !lat_ok(flow->q) || nb < 0 || ns < 0 || steps < 0) return -1;
  const size_t N = (size_t)flow->nx * (size_t)flow->ny * (size_t)flow->nz;
  REAL* A = pf;
  REAL* B = sf;
  REAL* SA[4];
  REAL* SB[4];
  if (ns > 4 || nb > 4) return -1;
  for (int k = 0; k < ns; ++k) {
    SA[k] = ps[k];
    SB[k] = ss[k];
  }
  for (int s = 0; s < steps; ++s) {
    for (size_t id = 0; id < N; ++id) flow->sforce[id] = 0.0;
    for (int k = 0; k < nb; ++k) {
      const REAL gb = (REAL)bu[k].gbeta, r0 = (REAL)bu[k].rho_init;
      const double* rs = bu[k].scalar_rho;
      for (size_t id = 0; id < N; ++id)
        if (flow->flags[id] & bu[k].mask) flow->sforce[id] = (double)((REAL)flow->sforce[id] + gb * ((REAL)rs[id] - r0));
    }
    FN(step)(flow, A, B);
    REAL* t = A;
    A = B;
    B = t;
    for (int k = 0; k < ns; ++k) {
      FN(step)(scal[k], SA[k], SB[k]);
      t = SA[k];
      SA[k] = SB[k];
      SB[k] = t;
    }
  }
  if (A != pf) memcpy(pf, A, sizeof(REAL) * (size_t)flow->q * N);
  for (int k = 0; k < ns; ++k)
    if (SA[k] != ps[k]) memcpy(ps[k], SA[k], sizeof(REAL) * (size_t)scal[k]->q * N);
  return 0;
}

void FN(rhou_impl)(const flbo_lattice* L, const REAL* pops, uint8_t mask, REAL* rhou) {
  const int q = L->q;
  const int d = lat_d(q);
  const size_t N = (size_t)L->nx * (size_t)L->ny * (size_t)L->nz;
#pragma omp parallel for schedule(static)
  for (size_t id = 0; id < N; ++id) {
    REAL c[19], rho = 0, u[3] = {0, 0, 0};
    if (L->flags[id] & mask) {
      for (int i = 0; i < q; ++i) c[i] = pops[(size_t)i * N + id];
      FN(rhou_cell)(q, c, &rho, u);
    }
    rhou[id] = rho;
    for (int a = 0; a < d; ++a) rhou[(size_t)(a + 1) * N + id] = u[a];
  }
}

void FN(init_feq_impl)(const flbo_lattice* L, const REAL* rhou, REAL* pops) {
  const int q = L->q;
  const int d = lat_d(q);
  const size_t N = (size_t)L->nx * (size_t)L->ny * (size_t)L->nz;
  REAL w[19];
  FN(weights)(q, w);
#pragma omp parallel for schedule(static)
  for (size_t id = 0; id < N; ++id) {
    REAL u[3] = {0, 0, 0}, feq[19];
    for (int a = 0; a < d; ++a) u[a] = rhou[(size_t)(a + 1) * N + id];
    FN(feq_cell)(q, w, rhou[id], u, feq);
    for (int i = 0; i < q; ++i) pops[(size_t)i * N + id] = feq[i];
  }
}


/* ---- two-level refinement (see flb_oracle.h) ---- */

/* Equilibrium<T, LatSet>::Order2 / SecondOrder (equilibrium.h:49-53, 69-75):
 * uc = u . c_k and u2 = |u|^2 summed from 0 in component order,
 * w rho (1 + 3 uc + uc uc 0.5 9 - 3 u2 0.5) evaluated left to right. */
static void FN(feq_order2)(int q, const REAL* w, REAL rho, const REAL* u, REAL* e) {
  const int d = lat_d(q);
  REAL u2 = 0;
  for (int a = 0; a < d; ++a) u2 += u[a] * u[a];
  for (int k = 0; k < q; ++k) {
    const int* ck = &lat_c(q)[k][0];
    REAL uc = 0;
    for (int a = 0; a < d; ++a) uc += u[a] * (REAL)ck[a];
    e[k] = w[k] * rho * ((REAL)1 + (REAL)3 * uc + uc * uc * (REAL)0.5 * (REAL)9 - (REAL)3 * u2 * (REAL)0.5);
  }
}

static REAL FN(field_rho)(const flbo_lattice* L, size_t id) {
  return L->rho_field ? (REAL)L->rho_field[id] : (REAL)1;
}
static REAL FN(field_u)(const flbo_lattice* L, int a, size_t id, size_t N) {
  return L->u_field ? (REAL)L->u_field[(size_t)a * N + id] : (REAL)0;
}

void CAT(flbo_refine_average, SFX)(const flbo_lattice* C, REAL* popsC, const flbo_lattice* F, const REAL* popsF,
                                   const int* origin) {
  const int q = C->q, d = lat_d(q), ns = 1 << d;
  const size_t NC = (size_t)C->nx * C->ny * C->nz, NF = (size_t)F->nx * F->ny * F->nz;
  const int m[3] = {(F->nx - 4) / 2, (F->ny - 4) / 2, d == 3 ? (F->nz - 4) / 2 : 1};
  const REAL omega_f = (REAL)F->omega, scale = d == 3 ? (REAL)0.125 : (REAL)0.25;
  REAL w[19];
  FN(weights)(q, w);
  for (int z = 0; z < m[2]; ++z)
    for (int y = 0; y < m[1]; ++y)
      for (int x = 0; x < m[0]; ++x) {
        if (!(x == 0 || x == m[0] - 1 || y == 0 || y == m[1] - 1 || (d == 3 && (z == 0 || z == m[2] - 1))))
          continue;
        size_t src[8];
        for (int j = 0; j < ns; ++j) {
          const int fx = 2 + 2 * x + (j & 1), fy = 2 + 2 * y + ((j >> 1) & 1);
          const int fz = d == 3 ? 2 + 2 * z + ((j >> 2) & 1) : 0;
          src[j] = (size_t)fx + (size_t)F->nx * ((size_t)fy + (size_t)F->ny * fz);
        }
        /* getAverage: (a0 + a1 + ... + a(n-1)) * 1/n */
        REAL rho = FN(field_rho)(F, src[0]), u[3] = {0, 0, 0};
        for (int j = 1; j < ns; ++j) rho = rho + FN(field_rho)(F, src[j]);
        rho = rho * scale;
        for (int a = 0; a < d; ++a) {
          REAL s = FN(field_u)(F, a, src[0], NF);
          for (int j = 1; j < ns; ++j) s = s + FN(field_u)(F, a, src[j], NF);
          u[a] = s * scale;
        }
        REAL e[19];
        FN(feq_order2)(q, w, rho, u, e);
        const size_t cid = (size_t)(origin[0] + x) +
                           (size_t)C->nx * ((size_t)(origin[1] + y) + (size_t)C->ny * (d == 3 ? origin[2] + z : 0));
        for (int k = 0; k < q; ++k) {
          const REAL* pk = popsF + (size_t)k * NF;
          REAL s = pk[src[0]];
          for (int j = 1; j < ns; ++j) s = s + pk[src[j]];
          const REAL p = s * scale;
          popsC[(size_t)k * NC + cid] = p + (REAL)0.5 * omega_f * (p - e[k]);
        }
      }
}

void CAT(flbo_refine_interpolate, SFX)(const flbo_lattice* C, const REAL* popsC, const flbo_lattice* F,
                                       REAL* popsF, const int* origin) {
  const int q = C->q, d = lat_d(q), ns = 1 << d;
  const size_t NC = (size_t)C->nx * C->ny * C->nz, NF = (size_t)F->nx * F->ny * F->nz;
  const int fn[3] = {F->nx, F->ny, d == 3 ? F->nz : 1};
  const REAL omega_c = (REAL)C->omega;
  REAL w[19];
  FN(weights)(q, w);
  for (int z = 0; z < fn[2]; ++z)
    for (int y = 0; y < fn[1]; ++y)
      for (int x = 0; x < fn[0]; ++x) {
        const int f3[3] = {x, y, z};
        int ghost = 0;
        for (int a = 0; a < d; ++a) ghost |= f3[a] < 2 || f3[a] >= fn[a] - 2;
        if (!ghost) continue;
        /* coarse unit u = origin + floor((f - 2) / 2); sub-cell bit (f - 2) mod 2;
           stencil corner u - 1 + bit (AddtoSharedIntpComm's Cid0 and its shifts) */
        int corner[3] = {0, 0, 0}, sub = 0;
        for (int a = 0; a < d; ++a) {
          const int rel = f3[a] - 2;
          const int fl = rel >= 0 ? rel / 2 : -((1 - rel) / 2);
          const int bit = rel - 2 * fl;
          corner[a] = origin[a] + fl - 1 + bit;
          sub |= bit << a;
        }
        size_t src[8];
        REAL wt[8];
        for (int j = 0; j < ns; ++j) {
          const int cx = corner[0] + (j & 1), cy = corner[1] + ((j >> 1) & 1);
          const int cz = d == 3 ? corner[2] + ((j >> 2) & 1) : 0;
          src[j] = (size_t)cx + (size_t)C->nx * ((size_t)cy + (size_t)C->ny * cz);
          /* getIntpWeight: 3/4 along each axis where the stencil cell is the
             nearer one (its bit differs from the sub-cell's), else 1/4 */
          REAL wj = 1;
          for (int a = 0; a < d; ++a) wj *= (((sub >> a) & 1) != ((j >> a) & 1)) ? (REAL)0.75 : (REAL)0.25;
          wt[j] = wj;
        }
        /* getInterpolation: a0 w0 + a1 w1 + ... left to right */
        REAL rho = FN(field_rho)(C, src[0]) * wt[0], u[3] = {0, 0, 0};
        for (int j = 1; j < ns; ++j) rho = rho + FN(field_rho)(C, src[j]) * wt[j];
        for (int a = 0; a < d; ++a) {
          REAL s = FN(field_u)(C, a, src[0], NC) * wt[0];
          for (int j = 1; j < ns; ++j) s = s + FN(field_u)(C, a, src[j], NC) * wt[j];
          u[a] = s;
        }
        REAL e[19];
        FN(feq_order2)(q, w, rho, u, e);
        const size_t fid = (size_t)x + (size_t)F->nx * ((size_t)y + (size_t)F->ny * z);
        for (int k = 0; k < q; ++k) {
          const REAL* pk = popsC + (size_t)k * NC;
          REAL s = pk[src[0]] * wt[0];
          for (int j = 1; j < ns; ++j) s = s + pk[src[j]] * wt[j];
          popsF[(size_t)k * NF + fid] = s - (REAL)0.25 * omega_c * (s - e[k]);
        }
      }
}

#undef FN
#undef CAT
#undef CAT_
