/* vfg_petsc.h -- PETSc adapter for libvfg: the reference's u- and
 * V-subproblem SNES objects (VFSolversInitialize, VFCommon.c:1012-1089) with
 * the GPU operators in place of the assembled Jacobians: a MATSHELL for each
 * MatMult, a PCSHELL for the u-KSP's multigrid, and the SNES function and
 * Jacobian callbacks (VF_UResidual / VF_UIJacobian, VFMech.c:1456, 1897;
 * VF_VResidual / VF_VIJacobian, VFMech.c:2687, 2883).
 *
 * Built by a PETSc host program, not by this tree (PETSc is no dependency of
 * libvfg):  $(CC) -c vfg_petsc.c $(PETSC_CC_INCLUDES) -I<repo>/include,
 * linked with -lvfg; PETSc 3.19 or later, real double scalars.  One MPI
 * rank per GPU whose DMDAs own the whole grid, so a global Vec is in the
 * natural order libvfg reads (the adapter checks it).
 */
#ifndef VFG_PETSC_H
#define VFG_PETSC_H
#include <petscsnes.h>
#include "vfg.h"

typedef struct {
  vfg_handle *h;  /* vfg_create + vfg_set_bc, as for the reference's VFCtx */
  /* the fields the operators are taken at (the reference's VFFields): V,
   * VIrrev, theta, thetaRef, pressure on daScal; E_elem one value per cell
   * (a sequential Vec) or NULL for the uniform E of the properties.  Changed
   * Vecs are pushed to the handle before each callback (PetscObjectState). */
  Vec V, VIrrev, theta, thetaRef, pressure, E_elem;
  Vec BCU; /* Dirichlet values of U (VecSetFromBC(BCU) times the load factor) */
  Vec U;   /* the displacement the V-subproblem is taken at */
  /* state of each Vec when last pushed: V VIrrev theta thetaRef pressure E_elem U */
  PetscObjectState st[7];
  PetscBool pushed[7];
} VFGPetscCtx;

/* zeroes g and sets g->h; the caller then sets the Vecs it has */
PetscErrorCode VFGPetscCtxInit(VFGPetscCtx *g, vfg_handle *h);
/* VFSolversInitialize with the shells: prefixes U_ and V_, SNESNEWTONLS with
 * the L2 line search and KSPCG for u (PCSHELL V-cycle unless -U_pc_type is
 * given), SNESVINEWTONRSLS on [0, 1] with KSPCG and PCJACOBI for V (the
 * reference's PCBJACOBI needs the assembled matrix). */
PetscErrorCode VFGPetscSolversInitialize(VFGPetscCtx *g, DM daVect, DM daScal, SNES *snesU, SNES *snesV);
PetscErrorCode VFGPetscSolversFinalize(SNES *snesU, SNES *snesV);
/* VF_StepU / VF_StepV (VFMech.c:1436-1451, 3057-3082) on the shells */
PetscErrorCode VFGPetscStepU(VFGPetscCtx *g, SNES snesU, Vec U);
PetscErrorCode VFGPetscStepV(VFGPetscCtx *g, SNES snesV, Vec V);
/* the same steps as whole GPU solves (vfg_u_solve / vfg_v_solve: multigrid
 * FP64 CG and the active-set V-solve, no PETSc solver); ub = VIrrev for AT1,
 * NULL (= 1) for AT2 */
PetscErrorCode VFGPetscSolveU(VFGPetscCtx *g, Vec U, const vfg_solve_opts *o, vfg_solve_info *info);
PetscErrorCode VFGPetscSolveV(VFGPetscCtx *g, Vec U, Vec V, Vec ub, const vfg_solve_opts *o, vfg_solve_info *info);

/* the callbacks, for hosts that build their own SNES (user = VFGPetscCtx *) */
PetscErrorCode VFGPetsc_UResidual(SNES snes, Vec U, Vec F, void *user);
PetscErrorCode VFGPetsc_UIJacobian(SNES snes, Vec U, Mat J, Mat P, void *user);
PetscErrorCode VFGPetsc_VResidual(SNES snes, Vec V, Vec F, void *user);
PetscErrorCode VFGPetsc_VIJacobian(SNES snes, Vec V, Mat J, Mat P, void *user);
/* the shells (context = VFGPetscCtx *) */
PetscErrorCode VFGPetscCreateUMat(VFGPetscCtx *g, DM daVect, Mat *J);
PetscErrorCode VFGPetscCreateVMat(VFGPetscCtx *g, DM daScal, Mat *J);
PetscErrorCode VFGPetscSetUPC(VFGPetscCtx *g, KSP ksp);
#endif
