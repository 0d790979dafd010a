/*
 * heat_oracle_impl.h -- type-generic body of the heat oracle, included by
 * heat_oracle.c once with RT=float (reference FP32 arithmetic) and once with
 * RT=double.  TEST INFRASTRUCTURE ONLY.  Loop orders and expression shapes
 * follow the reference's CPU branches cited per function.
 */
#define HCAT2(a, b) a##_##b
#define HCAT(a, b) HCAT2(a, b)
#define FN(name) HCAT(name, SFX)
#define CC(a, b, d) ((RT)c[(a) + (b) * nM + (d) * nM * nM])

/* substep1a, finiteElementHeatPropagator.c:288-312 */
static void FN(sub1a)(const heat_ctx *H, const float *c, const float *Q, RT dt, const RT *T1, RT *T2) {
  const long nx = H->nx, ny = H->ny, nz = H->nz, nM = H->nM, nxy = nx * ny;
  const unsigned char *M = H->M;
#pragma omp parallel for num_threads(H->nthreads) if (H->nthreads > 1)
  for (long iz = 0; iz < nz; iz++)
    for (long iy = 0; iy < ny; iy++)
      for (long ix = 0; ix < nx; ix++) {
        const long i = ix + iy * nx + iz * nxy;
        RT dTdt = 0;
        if (!frozen(H, ix, iy, iz)) {
          dTdt = H->lightsOn ? (RT)Q[i] : 0;
          if (ix != 0) dTdt += (T1[i - 1] - T1[i]) * CC(M[i], M[i - 1], 0) / 2;
          if (ix != nx - 1) dTdt += (T1[i + 1] - T1[i]) * CC(M[i], M[i + 1], 0) / 2;
          if (iy != 0) dTdt += (T1[i - nx] - T1[i]) * CC(M[i], M[i - nx], 1);
          if (iy != ny - 1) dTdt += (T1[i + nx] - T1[i]) * CC(M[i], M[i + nx], 1);
          if (iz != 0) dTdt += (T1[i - nxy] - T1[i]) * CC(M[i], M[i - nxy], 2);
          if (iz != nz - 1) dTdt += (T1[i + nxy] - T1[i]) * CC(M[i], M[i + nxy], 2);
        }
        T2[i] = T1[i] + dt * dTdt;
      }
}

/* Thomas solve of (I - dt/2 L_d) x = rhs along every line of axis d, in
 * place in T2 (substep1b :352-389, substep2b :486-521, substep3b :606-646).
 * The line of voxel i0 has n points at stride st; ib0/bst place the line's
 * forward-eliminated diagonal in the scratch b. */
static void FN(thomas_line)(const heat_ctx *H, const float *c, RT dt, int d, long i0, long st, long n, RT *T2, RT *b, long bst,
                            const long *pos /* ix,iy,iz of i0 */) {
  const long nM = H->nM;
  const unsigned char *M = H->M;
  long p[3] = {pos[0], pos[1], pos[2]};
  for (long k = 0; k < n; k++) {
    const long i = i0 + k * st;
    const long ib = k * bst;
    p[d] = k;
    b[ib] = 1;
    if (!frozen(H, p[0], p[1], p[2])) {
      if (k < n - 1) b[ib] += dt / 2 * CC(M[i], M[i + st], d);
      if (k > 0) {
        b[ib] += dt / 2 * CC(M[i], M[i - st], d);
        RT w = -dt / 2 * CC(M[i], M[i - st], d) / b[ib - bst];
        if (k > 1 || !H->sinked) b[ib] += w * dt / 2 * CC(M[i - st], M[i], d);
        T2[i] -= w * T2[i - st];
      }
    }
  }
  for (long k = n - 1; k >= 0; k--) {
    p[d] = k;
    if (!frozen(H, p[0], p[1], p[2])) {
      const long i = i0 + k * st;
      T2[i] = (T2[i] + (k == n - 1 ? 0 : dt / 2 * CC(M[i], M[i + st], d) * T2[i + st])) / b[k * bst];
    }
  }
}

/* substep2a (:429-448) for d = 1 and substep3a (:546-565) for d = 2 */
static void FN(sub_corr)(const heat_ctx *H, const float *c, RT dt, int d, const RT *T1, RT *T2) {
  const long nx = H->nx, ny = H->ny, nz = H->nz, nM = H->nM, nxy = nx * ny;
  const unsigned char *M = H->M;
  const long st = d == 1 ? nx : nxy;
  const long nd = d == 1 ? ny : nz;
#pragma omp parallel for num_threads(H->nthreads) if (H->nthreads > 1)
  for (long iz = 0; iz < nz; iz++)
    for (long iy = 0; iy < ny; iy++)
      for (long ix = 0; ix < nx; ix++) {
        const long i = ix + iy * nx + iz * nxy;
        if (frozen(H, ix, iy, iz)) continue;
        const long k = d == 1 ? iy : iz;
        RT dTdt = 0;
        if (k != 0) dTdt -= (T1[i - st] - T1[i]) * CC(M[i], M[i - st], d) / 2;
        if (k != nd - 1) dTdt -= (T1[i + st] - T1[i]) * CC(M[i], M[i + st], d) / 2;
        T2[i] = T2[i] + dt * dTdt;
      }
}

/* Arrhenius update and per-medium maximum, substep3b :639-642 */
static inline void FN(damage_and_max)(const heat_ctx *H, long i, RT dt, const RT *logA, const RT *E, RT Rgas, RT czero, const RT *T1,
                                      const RT *T2, RT *Omega, RT *maxT) {
  const unsigned char m = H->M[i];
  if (H->calcDamage && E[m]) Omega[i] += dt * (EXPF(logA[m] - E[m] / (Rgas * ((T2[i] + T1[i]) / 2 + czero))));
  if (maxT[m] < T2[i]) maxT[m] = T2[i];
}

/* forward-Euler step with the same neighbour order as substep1a */
static void FN(step_explicit)(const heat_ctx *H, const float *c, const float *Q, RT dt, const RT *logA, const RT *E, RT Rgas, RT czero,
                              const RT *T1, RT *T2, RT *Omega, RT *maxT) {
  const long nx = H->nx, ny = H->ny, nz = H->nz, nM = H->nM, nxy = nx * ny;
  const unsigned char *M = H->M;
#pragma omp parallel num_threads(H->nthreads) if (H->nthreads > 1)
  {
    RT mx[256];
    for (long m = 0; m < nM; m++) mx[m] = maxT[m];
#pragma omp for
    for (long iz = 0; iz < nz; iz++)
      for (long iy = 0; iy < ny; iy++)
        for (long ix = 0; ix < nx; ix++) {
          const long i = ix + iy * nx + iz * nxy;
          RT dTdt = 0;
          if (!frozen(H, ix, iy, iz)) {
            dTdt = H->lightsOn ? (RT)Q[i] : 0;
            if (ix != 0) dTdt += (T1[i - 1] - T1[i]) * CC(M[i], M[i - 1], 0);
            if (ix != nx - 1) dTdt += (T1[i + 1] - T1[i]) * CC(M[i], M[i + 1], 0);
            if (iy != 0) dTdt += (T1[i - nx] - T1[i]) * CC(M[i], M[i - nx], 1);
            if (iy != ny - 1) dTdt += (T1[i + nx] - T1[i]) * CC(M[i], M[i + nx], 1);
            if (iz != 0) dTdt += (T1[i - nxy] - T1[i]) * CC(M[i], M[i - nxy], 2);
            if (iz != nz - 1) dTdt += (T1[i + nxy] - T1[i]) * CC(M[i], M[i + nxy], 2);
          }
          T2[i] = T1[i] + dt * dTdt;
          FN(damage_and_max)(H, i, dt, logA, E, Rgas, czero, T1, T2, Omega, mx);
        }
    /* a maximum does not depend on the order it is taken in */
#pragma omp critical(orc_heat_max)
    for (long m = 0; m < nM; m++)
      if (maxT[m] < mx[m]) maxT[m] = mx[m];
  }
}

/* recordTempSensors, :652-663 */
static void FN(sensors)(const heat_ctx *H, const mcm_heat_input *in, const RT *T, double *out, long col) {
  const long nx = H->nx, nxy = H->nx * H->ny, nS = in->nSensors;
  for (long s = 0; s < nS; s++) {
    const long idx = (long)in->tempSensorCornerIdxs[s];
    const RT wx = (RT)in->tempSensorInterpWeights[s];
    const RT wy = (RT)in->tempSensorInterpWeights[s + nS];
    const RT wz = (RT)in->tempSensorInterpWeights[s + 2 * nS];
    out[s + col * nS] = (1 - wx) * (1 - wy) * (1 - wz) * T[idx] + (1 - wx) * (1 - wy) * wz * T[idx + nxy] + (1 - wx) * wy * (1 - wz) * T[idx + nx] +
                        (1 - wx) * wy * wz * T[idx + nx + nxy] + wx * (1 - wy) * (1 - wz) * T[idx + 1] + wx * (1 - wy) * wz * T[idx + 1 + nxy] +
                        wx * wy * (1 - wz) * T[idx + 1 + nx] + wx * wy * wz * T[idx + 1 + nx + nxy];
  }
}

/* mexFunction, :688-821; on H->nthreads threads (the values do not depend on it) */
static int FN(heat_run)(const heat_ctx *H, const mcm_heat_input *in, mcm_heat_output *out, int explicit_scheme) {
  const long nx = H->nx, ny = H->ny, nz = H->nz, nM = H->nM, L = nx * ny * nz, nxy = nx * ny;
  const long nt = in->steps;
  const int is_f32 = sizeof(RT) == sizeof(float);
  const RT dt = (RT)in->dt;
  const RT Rgas = is_f32 ? (RT)R_GAS_F : (RT)8.3144598;
  const RT czero = is_f32 ? (RT)CELSIUS_ZERO_F : (RT)273.15;
  RT *T1 = (RT *)malloc(L * sizeof(RT));
  RT *T2 = (RT *)malloc(L * sizeof(RT));
  RT *Om = (RT *)malloc((H->calcDamage ? L : 1) * sizeof(RT));
  /* one scratch row per thread, as long as the longest line */
  long bl = nx > ny ? nx : ny;
  if (bl < nz) bl = nz;
  RT *b = (RT *)malloc(bl * (H->nthreads > 1 ? H->nthreads : 1) * sizeof(RT));
  RT logA[256], E[256], maxT[256];
  if (!T1 || !T2 || !Om || !b || nM > 256) {
    free(T1); free(T2); free(Om); free(b);
    return MCM_ERR_OUT_OF_MEMORY;
  }
  for (long m = 0; m < nM; m++) {
    logA[m] = is_f32 ? (RT)(float)log(in->A[m]) : (RT)log(in->A[m]);
    E[m] = (RT)in->E[m];
    maxT[m] = (RT)-INFINITY;
  }
  for (long i = 0; i < L; i++) T1[i] = in->T_f64 ? (RT)in->T_f64[i] : (RT)in->T[i];
  if (H->calcDamage)
    for (long i = 0; i < L; i++) Om[i] = in->Omega_f64 ? (RT)in->Omega_f64[i] : (RT)in->Omega[i];
  if (in->nSensors && out->sensorTemps) FN(sensors)(H, in, T1, out->sensorTemps, 0);
  long n = 0;
  for (; n < nt; n++) {
    if (in->isInterruptPending && in->isInterruptPending()) break;
    if (explicit_scheme) {
      FN(step_explicit)(H, in->dTdtperdeltaT, in->dTdt_abs, dt, logA, E, Rgas, czero, T1, T2, Om, maxT);
    } else {
      const float *c = in->dTdtperdeltaT;
      FN(sub1a)(H, c, in->dTdt_abs, dt, T1, T2);
      /* every line solve is independent of the others, so lines may run on
       * several threads (each with its own scratch row of b) */
#pragma omp parallel for num_threads(H->nthreads) if (H->nthreads > 1)
      for (long iz = 0; iz < nz; iz++) {
        RT *bt = b + bl * omp_tid();
        for (long iy = 0; iy < ny; iy++) {
          long pos[3] = {0, iy, iz};
          FN(thomas_line)(H, c, dt, 0, iy * nx + iz * nxy, 1, nx, T2, bt, 1, pos);
        }
      }
      FN(sub_corr)(H, c, dt, 1, T1, T2);
#pragma omp parallel for num_threads(H->nthreads) if (H->nthreads > 1)
      for (long iz = 0; iz < nz; iz++) {
        RT *bt = b + bl * omp_tid();
        for (long ix = 0; ix < nx; ix++) {
          long pos[3] = {ix, 0, iz};
          FN(thomas_line)(H, c, dt, 1, ix + iz * nxy, nx, ny, T2, bt, 1, pos);
        }
      }
      FN(sub_corr)(H, c, dt, 2, T1, T2);
      /* the reference sweeps z with x innermost (:615-645); every z line is
       * independent, so line-at-a-time order gives the same values */
#pragma omp parallel for num_threads(H->nthreads) if (H->nthreads > 1)
      for (long iy = 0; iy < ny; iy++) {
        RT *bt = b + bl * omp_tid();
        for (long ix = 0; ix < nx; ix++) {
          long pos[3] = {ix, iy, 0};
          FN(thomas_line)(H, c, dt, 2, ix + iy * nx, nxy, nz, T2, bt, 1, pos);
        }
      }
#pragma omp parallel num_threads(H->nthreads) if (H->nthreads > 1)
      {
        RT mx[256];
        for (long m = 0; m < nM; m++) mx[m] = maxT[m];
#pragma omp for
        for (long iy = 0; iy < ny; iy++)
          for (long iz = nz - 1; iz >= 0; iz--)
            for (long ix = 0; ix < nx; ix++) FN(damage_and_max)(H, ix + iy * nx + iz * nxy, dt, logA, E, Rgas, czero, T1, T2, Om, mx);
#pragma omp critical(orc_heat_max)
        for (long m = 0; m < nM; m++)
          if (maxT[m] < mx[m]) maxT[m] = mx[m];
      }
    }
    RT *tmp = T1;
    T1 = T2;
    T2 = tmp;
    if (in->nSensors && out->sensorTemps) FN(sensors)(H, in, T1, out->sensorTemps, n + 1);
  }
  out->stepsDone = n;
  /* the reference leaves plhs[0] zero-filled when steps == 0 (:703-705, the
   * step loop never writes Tfinal); restated as is */
  if (out->T)
    for (long i = 0; i < L; i++) out->T[i] = nt == 0 ? 0.0f : (float)T1[i];
  if (out->T_f64)
    for (long i = 0; i < L; i++) out->T_f64[i] = (double)T1[i];
  if (out->Omega) {
    if (H->calcDamage)
      for (long i = 0; i < L; i++) out->Omega[i] = (float)Om[i];
    else
      out->Omega[0] = NAN;
  }
  if (out->Omega_f64 && H->calcDamage)
    for (long i = 0; i < L; i++) out->Omega_f64[i] = (double)Om[i];
  if (out->maxMediaTemps)
    for (long m = 0; m < nM; m++) out->maxMediaTemps[m] = (float)maxT[m];
  free(T1); free(T2); free(Om); free(b);
  return MCM_OK;
}

#undef HCAT2
#undef HCAT
#undef FN
#undef CC
