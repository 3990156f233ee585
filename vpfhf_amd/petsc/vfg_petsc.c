/* vfg_petsc.c -- the PETSc side of the libvfg drop-in: shells and SNES
 * callbacks in place of the reference's assembled u- and V-Jacobians
 * (VFSolversInitialize, VFCommon.c:1012-1089).  Each callback moves the Vec
 * arrays across the C-ABI of vfg.h (host arrays; with VECHIP Vecs the same
 * calls take VecHIPGetArray pointers and VFG_DEVICE). */
#include <petscdmda.h>
#include "vfg_petsc.h"

#if defined(PETSC_USE_COMPLEX) || !defined(PETSC_USE_REAL_DOUBLE)
#error "vfg_petsc.c needs real double-precision PetscScalar"
#endif

#define VFG_CALL(g, call)                                                                   \
  do {                                                                                      \
    if ((call) != VFG_OK) SETERRQ(PETSC_COMM_SELF, PETSC_ERR_LIB, "libvfg: %s", vfg_last_error((g)->h)); \
  } while (0)

enum { F_V, F_VIRREV, F_THETA, F_THETAREF, F_PRESSURE, F_E, F_U };

PetscErrorCode VFGPetscCtxInit(VFGPetscCtx *g, vfg_handle *h)
{
  PetscFunctionBegin;
  PetscCall(PetscMemzero(g, sizeof(*g)));
  g->h = h;
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* The DMDA must hold the whole grid on this rank: its global Vecs are then in
 * the natural order (dof fastest, then x, y, z) of the library's arrays. */
static PetscErrorCode check_whole_grid(VFGPetscCtx *g, DM da, PetscInt dof)
{
  PetscInt M, N, P, d, xs, ys, zs, xm, ym, zm;

  PetscFunctionBegin;
  PetscCall(DMDAGetInfo(da, NULL, &M, &N, &P, NULL, NULL, NULL, &d, NULL, NULL, NULL, NULL, NULL));
  PetscCall(DMDAGetCorners(da, &xs, &ys, &zs, &xm, &ym, &zm));
  PetscCheck(d == dof && xs == 0 && ys == 0 && zs == 0 && xm == M && ym == N && zm == P &&
                 (PetscInt64)M * N * P == vfg_num_nodes(g->h),
             PETSC_COMM_SELF, PETSC_ERR_ARG_SIZ,
             "vfg_petsc: the DMDA must hold the handle's whole grid (%" PetscInt_FMT " dof) on this rank", dof);
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* pushes field k to the handle if its Vec changed since the last push */
static PetscErrorCode push_field(VFGPetscCtx *g, int k, Vec v)
{
  PetscObjectState s;
  const PetscScalar *a;
  const double *f[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  int rc;

  PetscFunctionBegin;
  if (!v) PetscFunctionReturn(PETSC_SUCCESS);
  PetscCall(PetscObjectStateGet((PetscObject)v, &s));
  if (g->pushed[k] && s == g->st[k]) PetscFunctionReturn(PETSC_SUCCESS);
  PetscCall(VecGetArrayRead(v, &a));
  if (k == F_U) {
    rc = vfg_v_set_u(g->h, a, VFG_HOST);
  } else {
    f[k] = a;
    rc = vfg_set_fields(g->h, f[F_V], f[F_VIRREV], f[F_THETA], f[F_THETAREF], f[F_PRESSURE], f[F_E], VFG_HOST);
  }
  PetscCall(VecRestoreArrayRead(v, &a));
  VFG_CALL(g, rc);
  g->st[k] = s;
  g->pushed[k] = PETSC_TRUE;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode push_fields(VFGPetscCtx *g)
{
  PetscFunctionBegin;
  PetscCall(push_field(g, F_V, g->V));
  PetscCall(push_field(g, F_VIRREV, g->VIrrev));
  PetscCall(push_field(g, F_THETA, g->theta));
  PetscCall(push_field(g, F_THETAREF, g->thetaRef));
  PetscCall(push_field(g, F_PRESSURE, g->pressure));
  PetscCall(push_field(g, F_E, g->E_elem));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* ---------------------------------------------------------------- u side */
static PetscErrorCode UMult(Mat A, Vec x, Vec y) /* MatMult of JacU */
{
  VFGPetscCtx *g;
  const PetscScalar *xa;
  PetscScalar *ya;
  int rc;

  PetscFunctionBegin;
  PetscCall(MatShellGetContext(A, &g));
  PetscCall(VecGetArrayRead(x, &xa));
  PetscCall(VecGetArray(y, &ya));
  rc = vfg_u_apply(g->h, xa, ya, VFG_HOST);
  PetscCall(VecRestoreArray(y, &ya));
  PetscCall(VecRestoreArrayRead(x, &xa));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode UGetDiagonal(Mat A, Vec d)
{
  VFGPetscCtx *g;
  PetscScalar *da;
  int rc;

  PetscFunctionBegin;
  PetscCall(MatShellGetContext(A, &g));
  PetscCall(VecGetArray(d, &da));
  rc = vfg_u_diag(g->h, da, VFG_HOST);
  PetscCall(VecRestoreArray(d, &da));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode UPCApply(PC pc, Vec r, Vec z) /* one multigrid V-cycle */
{
  VFGPetscCtx *g;
  const PetscScalar *ra;
  PetscScalar *za;
  int rc;

  PetscFunctionBegin;
  PetscCall(PCShellGetContext(pc, &g));
  PetscCall(VecGetArrayRead(r, &ra));
  PetscCall(VecGetArray(z, &za));
  rc = vfg_u_pc_apply(g->h, ra, za, VFG_HOST);
  PetscCall(VecRestoreArray(z, &za));
  PetscCall(VecRestoreArrayRead(r, &ra));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode VFGPetscCreateUMat(VFGPetscCtx *g, DM daVect, Mat *J)
{
  PetscInt n = 3 * (PetscInt)vfg_num_nodes(g->h);

  PetscFunctionBegin;
  PetscCall(check_whole_grid(g, daVect, 3));
  PetscCall(MatCreateShell(PetscObjectComm((PetscObject)daVect), n, n, n, n, g, J));
  PetscCall(MatShellSetOperation(*J, MATOP_MULT, (void (*)(void))UMult));
  PetscCall(MatShellSetOperation(*J, MATOP_GET_DIAGONAL, (void (*)(void))UGetDiagonal));
  PetscCall(MatSetOption(*J, MAT_SYMMETRIC, PETSC_TRUE));
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode VFGPetscSetUPC(VFGPetscCtx *g, KSP ksp)
{
  PC pc;

  PetscFunctionBegin;
  PetscCall(KSPGetPC(ksp, &pc));
  PetscCall(PCSetType(pc, PCSHELL));
  PetscCall(PCShellSetContext(pc, g));
  PetscCall(PCShellSetApply(pc, UPCApply));
  PetscCall(PCShellSetName(pc, "vfg multigrid V-cycle"));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* VF_UResidual (VFMech.c:1456-1887): F = K U - b, U - BCU on Dirichlet rows */
PetscErrorCode VFGPetsc_UResidual(SNES snes, Vec U, Vec F, void *user)
{
  VFGPetscCtx *g = (VFGPetscCtx *)user;
  const PetscScalar *ua, *ba;
  PetscScalar *fa;
  int rc;

  PetscFunctionBegin;
  (void)snes;
  PetscCall(push_fields(g));
  PetscCall(VecGetArrayRead(U, &ua));
  PetscCall(VecGetArrayRead(g->BCU, &ba));
  PetscCall(VecGetArray(F, &fa));
  rc = vfg_u_residual(g->h, ua, ba, fa, VFG_HOST);
  PetscCall(VecRestoreArray(F, &fa));
  PetscCall(VecRestoreArrayRead(g->BCU, &ba));
  PetscCall(VecRestoreArrayRead(U, &ua));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* VF_UIJacobian (VFMech.c:1897-2029): nothing to assemble; the shell reads
 * the current fields, and under -unilateral NOCOMPRESSION the state U */
PetscErrorCode VFGPetsc_UIJacobian(SNES snes, Vec U, Mat J, Mat P, void *user)
{
  VFGPetscCtx *g = (VFGPetscCtx *)user;
  const PetscScalar *ua;
  int rc;

  PetscFunctionBegin;
  (void)snes;
  PetscCall(push_fields(g));
  PetscCall(VecGetArrayRead(U, &ua));
  rc = vfg_u_set_state(g->h, ua, VFG_HOST);
  PetscCall(VecRestoreArrayRead(U, &ua));
  VFG_CALL(g, rc);
  PetscCall(MatAssemblyBegin(J, MAT_FINAL_ASSEMBLY));
  PetscCall(MatAssemblyEnd(J, MAT_FINAL_ASSEMBLY));
  if (P != J) {
    PetscCall(MatAssemblyBegin(P, MAT_FINAL_ASSEMBLY));
    PetscCall(MatAssemblyEnd(P, MAT_FINAL_ASSEMBLY));
  }
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* ---------------------------------------------------------------- V side */
static PetscErrorCode VMult(Mat A, Vec x, Vec y)
{
  VFGPetscCtx *g;
  const PetscScalar *xa;
  PetscScalar *ya;
  int rc;

  PetscFunctionBegin;
  PetscCall(MatShellGetContext(A, &g));
  PetscCall(VecGetArrayRead(x, &xa));
  PetscCall(VecGetArray(y, &ya));
  rc = vfg_v_apply(g->h, xa, ya, VFG_HOST);
  PetscCall(VecRestoreArray(y, &ya));
  PetscCall(VecRestoreArrayRead(x, &xa));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode VGetDiagonal(Mat A, Vec d)
{
  VFGPetscCtx *g;
  PetscScalar *da;
  int rc;

  PetscFunctionBegin;
  PetscCall(MatShellGetContext(A, &g));
  PetscCall(VecGetArray(d, &da));
  rc = vfg_v_diag(g->h, da, VFG_HOST);
  PetscCall(VecRestoreArray(d, &da));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* SNESVINEWTONRSLS solves on the inactive set I with J_II
 * (MatCreateSubMatrix); the shell's J_II applies J to x_I extended by zero
 * and keeps the rows of I. */
typedef struct {
  Mat full;
  IS is;
  Vec xf, yf;
} VSub;

static PetscErrorCode VSubMult(Mat A, Vec x, Vec y)
{
  VSub *s;

  PetscFunctionBegin;
  PetscCall(MatShellGetContext(A, &s));
  PetscCall(VecSet(s->xf, 0.));
  PetscCall(VecISCopy(s->xf, s->is, SCATTER_FORWARD, x));
  PetscCall(MatMult(s->full, s->xf, s->yf));
  PetscCall(VecISCopy(s->yf, s->is, SCATTER_REVERSE, y));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode VSubGetDiagonal(Mat A, Vec d)
{
  VSub *s;

  PetscFunctionBegin;
  PetscCall(MatShellGetContext(A, &s));
  PetscCall(MatGetDiagonal(s->full, s->yf));
  PetscCall(VecISCopy(s->yf, s->is, SCATTER_REVERSE, d));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode VSubDestroy(Mat A)
{
  VSub *s;

  PetscFunctionBegin;
  PetscCall(MatShellGetContext(A, &s));
  PetscCall(ISDestroy(&s->is));
  PetscCall(VecDestroy(&s->xf));
  PetscCall(VecDestroy(&s->yf));
  PetscCall(PetscFree(s));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode VCreateSubMatrix(Mat A, IS isrow, IS iscol, MatReuse reuse, Mat *B)
{
  VSub *s;
  PetscInt m;
  PetscBool same;

  PetscFunctionBegin;
  PetscCall(ISEqual(isrow, iscol, &same));
  PetscCheck(same, PETSC_COMM_SELF, PETSC_ERR_SUP, "vfg_petsc: V submatrices are taken on equal row and column sets");
  if (reuse == MAT_REUSE_MATRIX) { /* same sizes: the existing shell takes the new index set */
    PetscCall(MatShellGetContext(*B, &s));
    PetscCall(PetscObjectReference((PetscObject)isrow));
    PetscCall(ISDestroy(&s->is));
    s->is = isrow;
    PetscCall(PetscObjectStateIncrease((PetscObject)*B));
    PetscFunctionReturn(PETSC_SUCCESS);
  }
  PetscCall(PetscNew(&s));
  s->full = A;
  PetscCall(PetscObjectReference((PetscObject)isrow));
  s->is = isrow;
  PetscCall(MatCreateVecs(A, &s->xf, &s->yf));
  PetscCall(ISGetLocalSize(isrow, &m));
  PetscCall(MatCreateShell(PetscObjectComm((PetscObject)A), m, m, m, m, s, B));
  PetscCall(MatShellSetOperation(*B, MATOP_MULT, (void (*)(void))VSubMult));
  PetscCall(MatShellSetOperation(*B, MATOP_GET_DIAGONAL, (void (*)(void))VSubGetDiagonal));
  PetscCall(MatShellSetOperation(*B, MATOP_DESTROY, (void (*)(void))VSubDestroy));
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode VFGPetscCreateVMat(VFGPetscCtx *g, DM daScal, Mat *J)
{
  PetscInt n = (PetscInt)vfg_num_nodes(g->h);

  PetscFunctionBegin;
  PetscCall(check_whole_grid(g, daScal, 1));
  PetscCall(MatCreateShell(PetscObjectComm((PetscObject)daScal), n, n, n, n, g, J));
  PetscCall(MatShellSetOperation(*J, MATOP_MULT, (void (*)(void))VMult));
  PetscCall(MatShellSetOperation(*J, MATOP_GET_DIAGONAL, (void (*)(void))VGetDiagonal));
  PetscCall(MatShellSetOperation(*J, MATOP_CREATE_SUBMATRIX, (void (*)(void))VCreateSubMatrix));
  PetscCall(MatSetOption(*J, MAT_SYMMETRIC, PETSC_TRUE));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* VF_VResidual (VFMech.c:2687-2879) at g->U */
PetscErrorCode VFGPetsc_VResidual(SNES snes, Vec V, Vec F, void *user)
{
  VFGPetscCtx *g = (VFGPetscCtx *)user;
  const PetscScalar *va;
  PetscScalar *fa;
  int rc;

  PetscFunctionBegin;
  (void)snes;
  PetscCall(push_fields(g));
  PetscCall(push_field(g, F_U, g->U));
  PetscCall(VecGetArrayRead(V, &va));
  PetscCall(VecGetArray(F, &fa));
  rc = vfg_v_residual(g->h, va, fa, VFG_HOST);
  PetscCall(VecRestoreArray(F, &fa));
  PetscCall(VecRestoreArrayRead(V, &va));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* VF_VIJacobian (VFMech.c:2883-3029): J_V depends on U only; refreshed when
 * U changed */
PetscErrorCode VFGPetsc_VIJacobian(SNES snes, Vec V, Mat J, Mat P, void *user)
{
  VFGPetscCtx *g = (VFGPetscCtx *)user;

  PetscFunctionBegin;
  (void)snes;
  (void)V;
  PetscCall(push_fields(g));
  PetscCall(push_field(g, F_U, g->U));
  PetscCall(MatAssemblyBegin(J, MAT_FINAL_ASSEMBLY));
  PetscCall(MatAssemblyEnd(J, MAT_FINAL_ASSEMBLY));
  if (P != J) {
    PetscCall(MatAssemblyBegin(P, MAT_FINAL_ASSEMBLY));
    PetscCall(MatAssemblyEnd(P, MAT_FINAL_ASSEMBLY));
  }
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* ------------------------------------------------------- solver objects */
PetscErrorCode VFGPetscSolversInitialize(VFGPetscCtx *g, DM daVect, DM daScal, SNES *snesU, SNES *snesV)
{
  SNESLineSearch ls;
  KSP ksp;
  PC pc;
  Mat JU, JV;
  Vec rU, rV, lb, ub;
  MPI_Comm comm = PetscObjectComm((PetscObject)daVect);

  PetscFunctionBegin;
  /* u: VFCommon.c:1020-1047 */
  PetscCall(SNESCreate(comm, snesU));
  PetscCall(SNESSetDM(*snesU, daVect));
  PetscCall(SNESSetOptionsPrefix(*snesU, "U_"));
  PetscCall(SNESSetType(*snesU, SNESNEWTONLS));
  PetscCall(SNESSetTolerances(*snesU, 1.e-8, 1.e-8, PETSC_DEFAULT, PETSC_DEFAULT, PETSC_DEFAULT));
  PetscCall(SNESSetFromOptions(*snesU));
  PetscCall(SNESGetLineSearch(*snesU, &ls));
  PetscCall(SNESLineSearchSetType(ls, SNESLINESEARCHL2));
  PetscCall(DMCreateGlobalVector(daVect, &rU));
  PetscCall(VFGPetscCreateUMat(g, daVect, &JU));
  PetscCall(SNESSetFunction(*snesU, rU, VFGPetsc_UResidual, g));
  PetscCall(SNESSetJacobian(*snesU, JU, JU, VFGPetsc_UIJacobian, g));
  PetscCall(VecDestroy(&rU));
  PetscCall(MatDestroy(&JU));
  PetscCall(SNESGetKSP(*snesU, &ksp));
  PetscCall(KSPSetTolerances(ksp, 1.e-8, 1.e-10, PETSC_DEFAULT, PETSC_DEFAULT));
  PetscCall(KSPSetType(ksp, KSPCG));
  PetscCall(VFGPetscSetUPC(g, ksp)); /* -U_pc_type jacobi|none overrides it */
  PetscCall(KSPSetFromOptions(ksp));

  /* V: VFCommon.c:1052-1088 */
  PetscCall(SNESCreate(comm, snesV));
  PetscCall(SNESSetDM(*snesV, daScal));
  PetscCall(SNESSetOptionsPrefix(*snesV, "V_"));
  PetscCall(SNESSetType(*snesV, SNESVINEWTONRSLS));
  PetscCall(DMCreateGlobalVector(daScal, &lb));
  PetscCall(VecSet(lb, 0.0));
  PetscCall(DMCreateGlobalVector(daScal, &ub));
  PetscCall(VecSet(ub, 1.0));
  PetscCall(SNESVISetVariableBounds(*snesV, lb, ub));
  PetscCall(VecDestroy(&lb));
  PetscCall(VecDestroy(&ub));
  PetscCall(SNESSetTolerances(*snesV, 1.e-8, 1.e-8, PETSC_DEFAULT, PETSC_DEFAULT, PETSC_DEFAULT));
  PetscCall(SNESSetFromOptions(*snesV));
  PetscCall(DMCreateGlobalVector(daScal, &rV));
  PetscCall(VFGPetscCreateVMat(g, daScal, &JV));
  PetscCall(SNESSetFunction(*snesV, rV, VFGPetsc_VResidual, g));
  PetscCall(SNESSetJacobian(*snesV, JV, JV, VFGPetsc_VIJacobian, g));
  PetscCall(VecDestroy(&rV));
  PetscCall(MatDestroy(&JV));
  PetscCall(SNESGetKSP(*snesV, &ksp));
  PetscCall(KSPSetTolerances(ksp, 1.e-8, 1.e-10, PETSC_DEFAULT, PETSC_DEFAULT));
  PetscCall(KSPSetType(ksp, KSPCG));
  PetscCall(KSPGetPC(ksp, &pc));
  PetscCall(PCSetType(pc, PCJACOBI));
  PetscCall(KSPSetFromOptions(ksp));
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode VFGPetscSolversFinalize(SNES *snesU, SNES *snesV)
{
  PetscFunctionBegin;
  PetscCall(SNESDestroy(snesU));
  PetscCall(SNESDestroy(snesV));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* VF_StepU (VFMech.c:1436-1451) */
PetscErrorCode VFGPetscStepU(VFGPetscCtx *g, SNES snesU, Vec U)
{
  SNESConvergedReason reason;
  PetscInt its;

  PetscFunctionBegin;
  PetscCall(push_fields(g));
  PetscCall(SNESSolve(snesU, NULL, U));
  PetscCall(SNESGetConvergedReason(snesU, &reason));
  if (reason < 0) {
    PetscCall(PetscPrintf(PETSC_COMM_WORLD, "[ERROR] snesU diverged with reason %d\n", (int)reason));
  } else {
    PetscCall(SNESGetIterationNumber(snesU, &its));
    PetscCall(PetscPrintf(PETSC_COMM_WORLD, "      snesU converged in %d iterations %d.\n", (int)its, (int)reason));
  }
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* VF_StepV (VFMech.c:3057-3082); returns the negative reason on divergence
 * like the reference */
PetscErrorCode VFGPetscStepV(VFGPetscCtx *g, SNES snesV, Vec V)
{
  SNESConvergedReason reason;
  PetscInt its;
  PetscReal vmin, vmax;
  PetscErrorCode flg = PETSC_SUCCESS;

  PetscFunctionBegin;
  PetscCall(push_fields(g));
  PetscCall(push_field(g, F_U, g->U));
  PetscCall(SNESSolve(snesV, NULL, V));
  PetscCall(SNESGetConvergedReason(snesV, &reason));
  if (reason < 0) {
    PetscCall(PetscPrintf(PETSC_COMM_WORLD, "[ERROR] snesV diverged with reason %d\n", (int)reason));
    flg = (PetscErrorCode)reason;
  } else {
    PetscCall(SNESGetIterationNumber(snesV, &its));
    PetscCall(PetscPrintf(PETSC_COMM_WORLD, "      snesV converged in %d iterations %d.\n", (int)its, (int)reason));
  }
  PetscCall(VecMin(V, NULL, &vmin));
  PetscCall(VecMax(V, NULL, &vmax));
  PetscCall(PetscPrintf(PETSC_COMM_WORLD, "      V min / max:     %e %e\n", (double)vmin, (double)vmax));
  PetscFunctionReturn(flg);
}

PetscErrorCode VFGPetscSolveU(VFGPetscCtx *g, Vec U, const vfg_solve_opts *o, vfg_solve_info *info)
{
  PetscScalar *ua;
  const PetscScalar *ba;
  int rc;

  PetscFunctionBegin;
  PetscCall(push_fields(g));
  PetscCall(VecGetArray(U, &ua));
  PetscCall(VecGetArrayRead(g->BCU, &ba));
  rc = vfg_u_solve(g->h, ua, ba, o, info, VFG_HOST);
  PetscCall(VecRestoreArrayRead(g->BCU, &ba));
  PetscCall(VecRestoreArray(U, &ua));
  VFG_CALL(g, rc);
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode VFGPetscSolveV(VFGPetscCtx *g, Vec U, Vec V, Vec ub, const vfg_solve_opts *o, vfg_solve_info *info)
{
  const PetscScalar *ua, *uba = NULL;
  PetscScalar *va;
  int rc;

  PetscFunctionBegin;
  PetscCall(push_fields(g));
  PetscCall(VecGetArrayRead(U, &ua));
  PetscCall(VecGetArray(V, &va));
  if (ub) PetscCall(VecGetArrayRead(ub, &uba));
  rc = vfg_v_solve(g->h, ua, va, uba, o, info, VFG_HOST);
  if (ub) PetscCall(VecRestoreArrayRead(ub, &uba));
  PetscCall(VecRestoreArray(V, &va));
  PetscCall(VecRestoreArrayRead(U, &ua));
  VFG_CALL(g, rc); /* V's new state makes the next callback push it */
  PetscFunctionReturn(PETSC_SUCCESS);
}
