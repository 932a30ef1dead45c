/*
 * A minimal stand-in for the part of PETSc's petscvec.h that csrc/compat.c
 * uses when compiled with -DFVM_WITH_PETSC (PETSc 3.5 prototypes).  It exists
 * only so that a CPU test can compile the PETSc variant of the wrappers without
 * PETSc: it declares, it implements nothing, and nothing links against it.
 */
#ifndef FVM_TEST_PETSCVEC_SHIM_H
#define FVM_TEST_PETSCVEC_SHIM_H

typedef struct _p_Vec *Vec;
typedef double PetscScalar;
typedef int PetscErrorCode;

PetscErrorCode VecGetArray(Vec, PetscScalar **);
PetscErrorCode VecRestoreArray(Vec, PetscScalar **);
PetscErrorCode VecGetArrayRead(Vec, const PetscScalar **);
PetscErrorCode VecRestoreArrayRead(Vec, const PetscScalar **);

#endif
