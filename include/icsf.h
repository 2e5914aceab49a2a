/*
 * icsf.h - C ABI of the MI355X implicit coupled pseudo-time step.
 *
 * Drop-in boundary for ICSFoam's hot path: the convective/viscous flux
 * schemes, the blockFvMatrix/coupledMatrix Jacobian and the coupledMatrix
 * GMRES/LU-SGS solve, driven by the dbnsFoam pseudo-time loop.  Every entry
 * point takes plain pointers and sizes; host arrays are caller-owned, device
 * buffers are owned by the library and persist across pseudo-steps.
 *
 * Unknown order at this interface: [rho, rhoE, rhoUx, rhoUy, rhoUz] per cell,
 * vectors laid out [5][nCells] (coupledMatrix with nScal = 2, nVect = 1,
 * applications/solvers/dbnsFoam/outerLoop.H:53).  5x5 blocks are row-major in
 * that order.  Cell-wise vector fields (U) are [nCells][3].
 *
 * Return codes: 0 on success, negative on failure (icsf_last_error() gives
 * the text).  Failure codes mirror the reference's FatalError sites.
 * Reference paths are relative to the ICSFoam source tree.
 */
#ifndef ICSF_H
#define ICSF_H

#ifdef __cplusplus
extern "C" {
#endif

#define ICSF_OK 0
#define ICSF_ERR_ARG (-1)        /* invalid argument / size */
#define ICSF_ERR_KEYWORD (-2)    /* unknown scheme/solver keyword: newConvectiveFluxScheme.C:51-60,
                                    coupledMatrixSolver.C:53-61, coupledMatrixPreconditioner.C:49 */
#define ICSF_ERR_ZERO_DIAG (-3)  /* lusgs.C:119-123 ("All diagonals of coupledMatrix are zero."): a
                                    cell's LU-SGS coefficient 1/max|diag| below VSMALL, i.e. a diagonal
                                    entry above 1/VSMALL (as the reference tests it; an all-zero
                                    diagonal gives 1/SMALL there and passes) or not a number */
#define ICSF_ERR_NONFINITE (-4)  /* non-finite input value, residual or conserved state */
#define ICSF_ERR_HIP (-5)        /* HIP runtime failure */
#define ICSF_ERR_STATE (-6)      /* call-order error (e.g. solve before assemble) */
#define ICSF_ERR_NODEVICE (-7)   /* no HIP device available */

/* patch kinds (OpenFOAM polyPatch types) */
#define ICSF_PATCH 0
#define ICSF_WALL 1
#define ICSF_EMPTY 2
#define ICSF_CYCLIC 3 /* cyclic (rotational or translational) pair member */
/* boundary-condition kinds (OpenFOAM fvPatchField types) */
#define ICSF_BC_ZERO_GRADIENT 0
#define ICSF_BC_FIXED_VALUE 1
#define ICSF_BC_SLIP 2                          /* U only */
#define ICSF_BC_TOTAL_PRESSURE 3                /* p only: value p0 (psi thermo:psi, gamma of the thermo) */
#define ICSF_BC_PRESSURE_INLET_OUTLET_VELOCITY 4 /* U only: value tangentialVelocity (NULL: zero) */
#define ICSF_BC_TOTAL_TEMPERATURE 5             /* T only: value T0 (gamma of the thermo) */

typedef struct icsf_mesh icsf_mesh;
typedef struct icsf_case icsf_case;
typedef struct icsf_hb icsf_hb;

/* Text of the last error on the calling thread. */
const char* icsf_last_error(void);
/* ABI version (major*100 + minor). */
int icsf_version(void);

/* ---------------------------------------------------------------------
 * Host-side mesh preprocessing (no GPU use).  Integer work; results are
 * deterministic and reproducible bit-for-bit.
 * ------------------------------------------------------------------- */

/* Reverse Cuthill-McKee cell order (new -> old).  Replaces OpenFOAM
 * renumberMesh's CuthillMcKee renumbering built on
 * meshTools::bandCompression, which ICSFoam meshes are prepared with;
 * ICSFoam itself runs on file order (src/blockFvMatrix/blockFvMatrix.C:220-234
 * consumes the resulting lduAddressing). */
int icsf_rcm_order(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int* newToOld);

/* The same on a mesh with coupled (cyclic) patches: each cell's neighbour
 * list is its internal-face neighbours followed by the cells across its
 * coupled faces, given as nCoupled (coupledOwner, coupledNeighbour) pairs
 * in boundary-face order; OpenFOAM renumberMesh's CuthillMcKeeRenumber
 * builds this list with decompositionMethod::calcCellCells. */
int icsf_rcm_order_coupled(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int nCoupled,
                           const int* coupledOwner, const int* coupledNeighbour, int* newToOld);

/* Greedy cell colouring in index order; returns the colour of each cell.
 * No counterpart in the reference (sweeps there are sequential,
 * src/coupledMatrix/preconditioners/lusgs/lusgs.C:233-381); used to build a
 * numbering in which the sequential LU-SGS sweep has few dependency levels. */
int icsf_colour_cells(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int* colour,
                      int* nColours);

/* Stable partition of cells by colour (new -> old). */
int icsf_colour_major_order(int nCells, const int* colour, int nColours, int* newToOld);

/* Apply a cell permutation to a polyMesh: internal faces re-oriented so
 * owner < neighbour (flipped faces reversed), re-sorted upper-triangular;
 * boundary faces keep their order.  faceNewToOld[f] = old face, or
 * -(old+1) when the face was flipped.  faceVertsOut has the same length as
 * faceVerts; face offsets of the output follow the new face order. */
int icsf_renumber_mesh(int nCells, int nFaces, int nInternalFaces, const int* faceOffsets, const int* faceVerts,
                       const int* owner, const int* neighbour, const int* newToOld, int* faceVertsOut,
                       int* ownerOut, int* neighbourOut, int* faceNewToOld);

/* Topology of an nx*ny*nz hexahedral block in blockMesh order (the
 * structured-mesh generator of the synthetic cases; replaces running
 * OpenFOAM blockMesh on a one-block blockMeshDict).  sidePatch[6] assigns
 * the sides (-x, +x, -y, +y, -z, +z) to patches 0..nPatches-1.  Outputs:
 * faceVerts [nFaces][4], owner [nFaces], neighbour [nInternalFaces],
 * patchSize [nPatches]; internal faces first (upper-triangular), then the
 * patches in order.  Point index of (i,j,k) is i + (nx+1)*(j + (ny+1)*k). */
int icsf_block_mesh_hex(int nx, int ny, int nz, const int* sidePatch, int nPatches, int* faceVerts, int* owner,
                        int* neighbour, int* patchSize);

/* Dependency levels of the forward (lower) and backward (upper) LU-SGS
 * sweeps in the current numbering (lusgs.C:233-381 visits cells in index
 * order; cells of one level are independent). */
int icsf_sweep_levels(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int* fwdLevel,
                      int* bwdLevel, int* nFwd, int* nBwd);

/* OpenFOAM mesh geometry on the host (primitiveMesh face/cell centres,
 * areas, volumes; surfaceInterpolation weights, deltaCoeffs,
 * nonOrthDeltaCoeffs, nonOrthCorrectionVectors).  Any output may be NULL. */
int icsf_mesh_geometry(int nPoints, const double* points, int nFaces, const int* faceOffsets, const int* faceVerts,
                       int nInternalFaces, const int* owner, const int* neighbour, double* C, double* V, double* Sf,
                       double* Cf, double* weights, double* deltaCoeffs, double* nonOrthDeltaCoeffs,
                       double* corrVecs);

/* ---------------------------------------------------------------------
 * Device mesh (fvMesh replacement for the path).
 * ------------------------------------------------------------------- */

/* Geometry of a mesh with patches, cyclic ones included: interpolation
 * weights, deltaCoeffs, nonOrthDeltaCoeffs and correction vectors of every
 * face (cyclic faces: cyclicFvPatch::makeWeights and cyclicFvPatch::delta),
 * and for every boundary face the cell across it (-1 off cyclic patches).
 * Patch arguments as icsf_mesh_create.  Any output may be NULL. */
int icsf_mesh_geometry_patches(int nPoints, const double* points, int nFaces, const int* faceOffsets,
                               const int* faceVerts, int nInternalFaces, const int* owner, const int* neighbour,
                               int nPatches, const int* patchStart, const int* patchSize, const int* patchKind,
                               const int* patchNeighbour, const double* patchTransform, double* weights,
                               double* deltaCoeffs, double* nonOrthDeltaCoeffs, double* corrVecs,
                               int* cyclicCell);

/* Upload a polyMesh (owner < neighbour, upper-triangular internal faces,
 * patches tiling the boundary faces in order).  Computes the geometry,
 * the face-ordered cell->face pull lists, the sliced-ELL block layout and
 * the LU-SGS sweep levels.  Cyclic patches (patchKind ICSF_CYCLIC) name
 * their paired patch in patchNeighbour[p] and carry forwardT, the 3x3
 * row-major tensor mapping the paired side's vectors into this side's
 * frame, in patchTransform[9p..9p+8] (OpenFOAM cyclicPolyPatch /
 * cyclicTransform); face i of a cyclic patch is coupled to face i of its
 * pair.  Both arrays may be NULL when there is no cyclic patch. */
int icsf_mesh_create(int nPoints, const double* points, int nFaces, const int* faceOffsets, const int* faceVerts,
                     int nInternalFaces, const int* owner, const int* neighbour, int nPatches,
                     const int* patchStart, const int* patchSize, const int* patchKind, const int* patchNeighbour,
                     const double* patchTransform, icsf_mesh** out);
int icsf_mesh_destroy(icsf_mesh* mesh);
/* info[0..7] = nCells, nFaces, nInternalFaces, nFwdLevels, nBwdLevels,
 * nSlices, sellSlots, emptyDirection (-1 if 3-D) */
int icsf_mesh_info(const icsf_mesh* mesh, int* info);

/* ---------------------------------------------------------------------
 * Case: fields, thermophysical model, schemes and solution controls.
 * ------------------------------------------------------------------- */

/* Dictionaries in OpenFOAM syntax: thermophysicalProperties (hePsiThermo,
 * hConst, perfectGas, const transport; keywords R, gamma, mu, Pr), fvSchemes
 * (convectiveFluxScheme{fluxScheme AUSMPlusUp|ROE|HLLC; lowMachAusm;
 * entropyFixCoeff}, ddtSchemes, gradSchemes, interpolationSchemes
 * reconstruct(rho|U|T)) and fvSolution (flowSolver{solver GMRES|FGMRES;
 * <solver>{preconditioner LUSGS|Jacobi|none; nDirections; maxIter; minIter;
 * tolerance; relTol; precision double|mixed; innerDirections;
 * orthogonalisation MGS|CGS; fusedSweeps, graphCapture, autotune,
 * fusedCellSums true|false; jacobianStorage generators|blocks}},
 * pseudoTime{...}).
 * Replaces convectiveFluxScheme::New (newConvectiveFluxScheme.C:39-64),
 * viscousFluxScheme construction (viscousFluxScheme.H:105-113),
 * coupledMatrix::solver::New (coupledMatrixSolver.C:41-67) and
 * coupledMatrix::preconditioner::New (coupledMatrixPreconditioner.C:41-64). */
int icsf_case_create(icsf_mesh* mesh, const char* thermophysicalProperties, const char* fvSchemes,
                     const char* fvSolution, icsf_case** out);
int icsf_case_destroy(icsf_case* c);

/* Boundary conditions of p, U, T on one patch; values per patch face (U as
 * [size][3]): the fixed value, or p0 of totalPressure, T0 of
 * totalTemperature, tangentialVelocity of pressureInletOutletVelocity;
 * pointers may be NULL where a condition takes no value.  Replaces the
 * boundaryField entries of 0/p, 0/U, 0/T as the ICSF tutorials set them
 * (tutorials/circularArcBump/transonic/0/{p,U,T}).  The derived inflow
 * conditions follow OpenFOAM-v2112 (rho none, psi thermo:psi, gamma of the
 * thermophysical model, phi the convective mass flux of the last residual
 * evaluation): they are corrected after every state update from lagged
 * values (updateFields.H:90-105), and enter the Jacobian through their
 * valueInternalCoeffs (convectiveFluxScheme.C:47-120).  Non-finite values
 * are rejected (ICSF_ERR_NONFINITE), unknown kinds with ICSF_ERR_KEYWORD. */
int icsf_set_boundary(icsf_case* c, int patch, int pType, int UType, int TType, const double* pValue,
                      const double* UValue, const double* TValue);

/* Initial boundary values of the derived conditions (the `value` entries
 * of the field files, read by createFields.H:51-145); any pointer may be
 * NULL.  Defaults: p0 for totalPressure (OpenFOAM's default without a
 * value entry), the owner cell's U for pressureInletOutletVelocity, T0 for
 * totalTemperature.  Takes effect at the next icsf_set_fields. */
int icsf_set_boundary_values(icsf_case* c, int patch, const double* pValue, const double* UValue,
                             const double* TValue);

/* Initial state from rho, U, T (createFields.H: p = rho/psi, rhoU, rhoE). */
int icsf_set_fields(icsf_case* c, const double* rho, const double* U, const double* T);
/* Current state; W is the conserved vector [5][nCells]. Any may be NULL. */
int icsf_get_fields(icsf_case* c, double* rho, double* U, double* T, double* p, double* W);
/* Conserved state W [5][nCells] as the new case state through updateFields
 * with a zero increment (updateFields.H:7-105: primitive recovery, bounds,
 * boundary conditions corrected); needs icsf_set_fields first, which
 * initialises the boundary state. */
int icsf_set_state(icsf_case* c, const double* W);
/* Boundary values of every boundary face as of the last state update
 * (p.boundaryField(), U.boundaryField(), T.boundaryField(); [nFaces -
 * nInternalFaces], U as [..][3]; zero on empty and cyclic faces).  Any may
 * be NULL. */
int icsf_get_boundary_values(icsf_case* c, double* p, double* U, double* T);

/* Face fluxes of every face (phi [nFaces], phiUp [nFaces][3], phiEp
 * [nFaces]); replaces convectiveFluxScheme::calcFlux
 * (convectiveFluxScheme.H:219-224). */
int icsf_calc_flux(icsf_case* c, double* phi, double* phiUp, double* phiEp);

/* Residual source b = R*V ([5][nCells]) of residualsUpdate.H:1-183 and the
 * local rPseudoDeltaT of setCoAndDeltaT.H:57-141.  Either may be NULL. */
int icsf_residual(icsf_case* c, double* b, double* rPseudoDeltaT);

/* ---- residual adjoint (no counterpart in the reference) --------------
 * Vector-Jacobian product of the residual b(W, Sf, Cf) with the AUSM+up,
 * Roe or HLLC scheme, inviscid or viscous (the forward chain of
 * updateFields.H:7-104, ausmPlusUpFluxScheme.C:73-299,
 * roeFluxScheme.C:41-241,276-409 or hllcFluxScheme.C:76-221, and
 * residualsUpdate.H:1-43), cell
 * centres, volumes and weights fixed:
 *   gW  = (db/dW)^T  v   [5][nCells]
 *   gSf = (db/dSf)^T v   [nFaces][3]   (NULL: not returned)
 *   gCf = (db/dCf)^T v   [nFaces][3]   (NULL: not returned)
 * W ([5][nCells], conserved) becomes the case state through updateFields
 * with a zero increment; W NULL keeps the current state.  gSf and gCf are
 * built for inviscid residuals (viscous cases must pass NULL); cyclic
 * meshes, frame zones and the derived inflow conditions return
 * ICSF_ERR_ARG.  Host
 * pointers; returns when the results are written. */
int icsf_residual_vjp(icsf_case* c, const double* W, const double* v, double* gW, double* gSf, double* gCf);
/* Same with device pointers (for framework bindings); synchronous. */
int icsf_residual_vjp_device(icsf_case* c, const double* W, const double* v, double* gW, double* gSf,
                             double* gCf);
/* icsf_residual_vjp_device without any device allocation (no counterpart
 * in the reference;
 * for callers that cannot allocate: code under HIP graph capture, serving
 * loops): the adjoint's work buffers (about 750 bytes per cell and 240 per
 * face, 24 and 170 more with the viscous terms) live in one fixed 128 MiB
 * array per GPU, held until the process ends; cases needing more are
 * rejected with ICSF_ERR_ARG.  The case's device must be current.  Calls on
 * one GPU share the array and must not overlap; results are those of
 * icsf_residual_vjp_device. */
int icsf_residual_vjp_device_noalloc(icsf_case* c, const double* W, const double* v, double* gW, double* gSf,
                                     double* gCf);
/* Records of the points where the residual of icsf_residual_vjp is not
 * differentiable (no counterpart in the reference; for derivative checks):
 * W (device, [5][nCells], or NULL) becomes the state as in the VJP;
 * faceBranches [nFaces] gets each face's flux switches as bits - AUSM+up:
 * c* against the normal velocity on either side, the smaller c, subsonic
 * or supersonic split Mach and its sign on either side, the
 * pressure-diffusion clamp, the 0.2 damping, the upwind side (boundary
 * faces, whose two sides are one state: the smaller-c and upwind-side bits
 * 0); Roe: the signs of u, u + c, u - c, the entropy fix of each wave, the
 * larger of |u + c| and |u - c|, the sign of c^2's argument, the two
 * density clamps (boundary faces: 0, the dissipation there acting on a
 * zero jump) -, limiterElement [nCells][8] the
 * cell or boundary face (nCells + b) that is the max/min of each
 * reconstructed field's limiter (-1: none), limiterFace [nCells][8] the face
 * that set the limiter (-1: limiter 1); max/min candidates within 1e-12
 * relative of each other count as equal (the first is kept).  Two states
 * with equal records lie in one smooth piece of the residual.  Device
 * pointers; synchronous. */
int icsf_residual_records_device(icsf_case* c, const double* W, int* faceBranches, int* limiterElement,
                                 int* limiterFace);
/* b(W) of icsf_residual with device pointers: W ([5][nCells] or NULL for
 * the current state) becomes the case state, b receives R*V.  Synchronous. */
int icsf_residual_device(icsf_case* c, const double* W, double* b);

/* Assemble the coupled Jacobian for the current state (after
 * icsf_residual): convectiveFluxScheme::createConvectiveJacobian
 * (convectiveFluxScheme.C:537-546) and viscousFluxScheme::
 * createViscousJacobian (viscousFluxScheme.C:265-273) into the block
 * matrix (blockFvMatrix.C:212-326). */
int icsf_assemble_jacobian(icsf_case* c);

/* LDU view of the assembled matrix: diag [nCells][25], upper/lower
 * [nInternalFaces][25] (blockFvMatrix diag/upper/lower).  Any may be NULL. */
int icsf_get_jacobian(icsf_case* c, double* diag, double* upper, double* lower);
/* Interface blocks of the boundary faces [nFaces - nInternalFaces][25]
 * (blockFvMatrix interfacesUpper, blockFvMatrix.C:241-265 and :297-323):
 * the coupling block of each cyclic face, applied to the far cell's
 * vector with its momentum part transformed by forwardT; zero on the
 * other boundary faces. */
int icsf_get_interface_blocks(icsf_case* c, double* blocks);

/* Launch parameters in use (no counterpart in the reference): params[0..9]
 * = workgroup size of the Arnoldi-operator row kernels (SpMV, LU-SGS),
 * of the face-terms kernel, of the gradient and cell-residual kernels, of
 * the Jacobian rows; basis vectors per CGS projection pass; 1 if the
 * autotuner has run; then the register budget (0 the compiler's choice,
 * 1 at most 128 VGPRs for 4 waves per SIMD, 2 at most 168 for 3) of the
 * operator, face-terms,
 * gradient and Jacobian kernels.  fvSolution keyword autotune (default
 * true): the first pseudo-step times each candidate on the case's own data
 * and keeps the fastest; results are the same bits for every choice. */
int icsf_get_tuning(icsf_case* c, int* params);
/* Adds a rotating zone (MRF; replaces MRFCoupledZoneList / MRFProperties,
 * MRFCoupledZone.C:42-675, and the frame terms of the flux schemes and of
 * the Jacobian): the zone's cells, origin, axis (normalised here), omega
 * [rad/s] and the patches that do not rotate (nonRotatingPatches).  Faces of
 * the zone transport with U.n - um, um = (Omega ^ (Cf - origin)) . Sf/|Sf|;
 * the coupled system gets -(Omega ^ rho U) V and its derivative; the zone
 * cells' faces on the other non-empty, non-coupled patches rotate with it:
 * their U condition must be fixedValue and its value becomes
 * Omega ^ (Cf - origin).  Call after the boundary conditions and before the
 * fields.  The frame divergence puts a diagonal term on every off-diagonal
 * block; with jacobianStorage generators it is stored as one more value
 * per slot beside the ten generators.  Each call adds
 * the zone list's next entry: as MRFCoupledZoneList (MRFCoupledZoneList.C:
 * 55-142), every zone in call order assigns its faces' um, its cells' Omega
 * and its rotating faces' U, so a later zone overrides an earlier one where
 * they meet.  The residual VJP refuses such cases. */
int icsf_set_mrf(icsf_case* c, int nZoneCells, const int* zoneCells, const double* origin, const double* axis,
                 double omega, int nNonRotating, const int* nonRotatingPatches);
/* One translating zone (replaces MRFTranslatingZoneList /
 * MRFTranslatingProperties, MRFTranslatingZone.C; outerLoop.H:7,21): the
 * zone's cells, frameVelocity and nonMovingPatches.  The same face sets as
 * the rotating zone; its faces add frameVelocity . Sf/|Sf| to the face
 * velocity the schemes see, its moving patches (fixedValue U) take the
 * frame velocity (after the rotating zones'), no Coriolis term.  One per
 * case, beside any rotating zones; the same storage and VJP refusal
 * as icsf_set_mrf. */
int icsf_set_translating_frame(icsf_case* c, int nZoneCells, const int* zoneCells, const double* velocity,
                               int nNonMoving, const int* nonMovingPatches);
/* The same zones for every instance of a harmonic-balance case (the
 * instances share the mesh and the frame); after the instances' boundary
 * conditions, before their fields. */
int icsf_hb_set_mrf(icsf_hb* h, int nZoneCells, const int* zoneCells, const double* origin, const double* axis,
                    double omega, int nNonRotating, const int* nonRotatingPatches);
int icsf_hb_set_translating_frame(icsf_hb* h, int nZoneCells, const int* zoneCells, const double* velocity,
                                  int nNonMoving, const int* nonMovingPatches);
/* Time shift of one side of a phase-lagged cyclic pair
 * (phaseLagCyclicFvPatchField.C:289-326): Dl [nInstants][nInstants]
 * (row-major) = Re(E^-1 M E), E = pinv(E^-1) of the instants, M =
 * diag(1, e^{i n IBPA} for n = 1..nH, e^{-i n IBPA} at nT - n) over the
 * omega list of icsf_hb_operator.  The owner side of a pair takes the
 * patch's IBPA, the neighbour side minus the owner's IBPA
 * (phaseLagCyclicFvPatchFields.C:51-75). */
int icsf_hb_phase_lag_operator(int nFreq, const double* frequencies, const int* harmonics, int nInstants,
                               const double* instants, double ibpa, double* Dl);
/* Makes one side of a cyclic pair phase-lagged (replaces the
 * phaseLagCyclic patch type: phaseLagCyclicPolyPatch.C:79-90 IBPA,
 * HBZoneName, cylCoords; phaseLagCyclicFvPatchField.C:179-396): on its
 * faces the far-side values of the fields the reference lags (U and its
 * components, p, rho, gamma, c, sqrt(gamma/psi), E and H, as
 * phaseLagCyclicFvPatchField.C:351-358 names them) are sum_J Dl[I][J]
 * x_J(far cell) over the instances, in the residual, the local time step,
 * the Jacobian's face states and the bounding of the local time step;
 * gradients and the linear solver's interfaces stay the plain cyclic ones.
 * Call once per side, with that side's Dl; cylCoords (nonzero) combines
 * the velocities in cylindrical components about rotationAxis through
 * rotationCentre (the patch's), else the velocity components are
 * combined unrotated as the reference does (refused for viscous cases,
 * whose terms take grad(U) whole).  Set before the instances' fields or
 * after them; the next residual uses it. */
int icsf_hb_set_phase_lag(icsf_hb* h, int patch, const double* Dl, int cylCoords, const double* rotationAxis,
                          const double* rotationCentre);
/* Sets the launch parameters (the layout of icsf_get_tuning; params[5]
 * ignored) and marks the case tuned, so the autotuner does not run (no
 * counterpart in the reference; for pinning a measured choice).
 * Workgroup sizes 64, 128 or 256; projection passes 8, 16 or 32; budgets
 * as above (operator 0 or 1); anything else returns ICSF_ERR_ARG. */
int icsf_set_tuning(icsf_case* c, const int* params);
/* Replace the assembled matrix by given LDU blocks (same layout). */
int icsf_set_jacobian(icsf_case* c, const double* diag, const double* upper, const double* lower);

/* y = A x ([5][nCells]); coupledMatrix::matrixMul (coupledMatrix.C:66-123). */
int icsf_matrix_mul(icsf_case* c, const double* x, double* y);

/* r <- P^-1 r with the configured preconditioner; lusgs::precondition
 * (lusgs.C:220-382) / Jacobi::precondition (Jacobi.C:55-132). */
int icsf_precondition(icsf_case* c, double* r);

/* Solve A dW = b for the increment: coupledMatrix::solveForIncr
 * (coupledMatrix.C:297-385) -> gmres::solveDelta (gmres.C:772-1108).
 * W: current state [5][nCells] (NULL = the case state), b: source (NULL =
 * the last residual), dW out.  residuals[10] = sInit[2], vInit[3],
 * sFinal[2], vFinal[3] (residualsIO), nIter = cycles run. */
int icsf_solve_for_incr(icsf_case* c, const double* W, const double* b, double* dW, double* residuals,
                        int* nIter);

/* nSteps steady pseudo-time steps of dbnsFoam (dbnsFoam.C:110-124:
 * outerLoop.H + updateFields.H, SER of setCoAndDeltaT.H:3-37, bounding of
 * boundLocalTimeStep.H).  residuals: [nSteps][10] (may be NULL), nIter:
 * [nSteps] (may be NULL).  Residuals are read back only when requested.
 * Afterwards the state is the updated one: the residual and the matrix of
 * the last step refer to the state before its update, so the matrix calls
 * (icsf_solve_for_incr, icsf_matrix_mul, ...) need icsf_residual and
 * icsf_assemble_jacobian again (ICSF_ERR_STATE otherwise).  A conserved
 * state that turns non-finite during the steps is reported as
 * ICSF_ERR_NONFINITE after the last step (the device flags it in the state
 * update; graph replays do not stop early). */
int icsf_pseudo_steps(icsf_case* c, int nSteps, double* residuals, int* nIter);

/* The steady pseudo-time loop as dbnsFoam runs it (dbnsFoam.C:105-124 with
 * pseudotimeControl::loop, pseudotimeControl.C:72-101,159-241): before each
 * pseudo-step the criteria of fvSolution pseudoTime (pseudoTol,
 * pseudoTolRel -- both required, ICSF_ERR_KEYWORD otherwise --, nPseudoCorr,
 * nPseudoCorrMin) are evaluated on the initial residuals of the previous
 * step: absolute max(initial) < pseudoTol, or relative max(initial /
 * initial of step 1) < pseudoTolRel (from step 3 on); no check at step 1 or
 * at step nPseudoCorr.  When they hold, that step is the final one.  The
 * check runs on the device inside every (graph-replayed) step; the host
 * issues `batch` steps between reads of the end flag, and steps issued after
 * the end leave the state unchanged.  At most maxSteps steps; stepsRun and
 * converged (1 if the criteria ended the loop) may be NULL. */
int icsf_pseudo_time_loop(icsf_case* c, int maxSteps, int batch, int* stepsRun, int* converged);

/* Pseudo-Courant field [nCells] (pseudoCoField). */
int icsf_get_pseudo_co(icsf_case* c, double* co);

/* Wait for all work queued on the case's stream. */
int icsf_synchronize(icsf_case* c);

/* Kernel timing with HIP events recorded on the launch stream around each
 * launch (or launch group) of the selected kernel classes.
 * icsf_timing_enable(c, mask) resets and starts accumulating for classes
 * whose bit is set (mask 0 stops); icsf_timing_read returns, per class
 * k < n, total milliseconds and event-pair count.  Classes: 0 residual
 * (boundary + gradient + residual kernels; in a pseudo-step up to the face
 * terms), 1 Jacobian (in a pseudo-step with the residual's cell sums),
 * 2 Arnoldi operator (the block SpMV of each Arnoldi step, with the fused
 * LU-SGS sweeps where they apply), 3 preconditioner (all sweep levels of
 * one application), 4 orthogonalisation + Givens of one Arnoldi step,
 * 5 other (including the outer product of FGMRES). */
int icsf_timing_enable(icsf_case* c, int mask);
int icsf_timing_read(icsf_case* c, int n, double* ms, long long* count);

/* ---- OpenFOAM case files (host; no GPU use) --------------------------
 * ASCII files as OpenFOAM writes them; each may instead be present as
 * "<name>.gz" (gzip), which OpenFOAM reads transparently. */
typedef struct icsf_foam_mesh icsf_foam_mesh;
typedef struct icsf_foam_field icsf_foam_field;

/* Read constant/polyMesh (points, faces as faceList or faceCompactList,
 * owner, neighbour, boundary), as fvMesh construction reads it for
 * dbnsFoam (createDynamicFvMesh.H).  Patch types patch, wall, empty, cyclic,
 * phaseLagCyclic (kind ICSF_CYCLIC), symmetryPlane/symmetry (kind
 * ICSF_PATCH); cyclic pairs get their
 * neighbourPatch index and forwardT: identity for translational pairs,
 * for rotational ones the rotation about rotationAxis taking the paired
 * patch's mean normal to minus this patch's.  Sizes are checked (owner /
 * faces, patches tiling the boundary faces in order, vertex labels). */
int icsf_foam_mesh_read(const char* polyMeshDir, icsf_foam_mesh** out);
/* sizes[0..5] = nPoints, nFaces, nInternalFaces, nCells, nPatches, length
 * of the face-vertex list */
int icsf_foam_mesh_sizes(const icsf_foam_mesh* m, int* sizes);
/* The arrays icsf_mesh_create takes (any may be NULL); patchTransform
 * [nPatches][9], zero off cyclic patches. */
int icsf_foam_mesh_get(const icsf_foam_mesh* m, double* points, int* faceOffsets, int* faceVerts, int* owner,
                       int* neighbour, int* patchStart, int* patchSize, int* patchKind, int* patchNeighbour,
                       double* patchTransform);
int icsf_foam_mesh_patch_name(const icsf_foam_mesh* m, int patch, char* name, int len);
/* A phaseLagCyclic patch (read as a cyclic one; phaseLagCyclicPolyPatch.C:
 * 79-90): *phaseLag 1 and its IBPA, cylCoords, HBZoneName, rotationAxis
 * and rotationCentre; 0 for any other patch (outputs may be NULL). */
int icsf_foam_mesh_patch_phase_lag(const icsf_foam_mesh* m, int patch, int* phaseLag, double* ibpa, int* cylCoords,
                                   double* rotationAxis, double* rotationCentre, char* hbZoneName, int len);
int icsf_foam_mesh_free(icsf_foam_mesh* m);
/* Write a polyMesh directory (the five files, ASCII) from the arrays of
 * icsf_mesh_create; cyclic patches get neighbourPatch and a translational
 * or rotational transform. */
int icsf_foam_mesh_write(const char* polyMeshDir, int nPoints, const double* points, int nFaces,
                         const int* faceOffsets, const int* faceVerts, int nInternalFaces, const int* owner,
                         const int* neighbour, int nPatches, const char* const* patchNames, const int* patchStart,
                         const int* patchSize, const int* patchKind, const int* patchNeighbour,
                         const double* patchTransform);

/* Read a volScalarField or volVectorField file of nCells cells (the fields
 * createFields.H:51-145 reads: p, U, T): internalField given uniform or
 * nonuniform; boundaryField entries queried by patch name. */
int icsf_foam_field_read(const char* path, int nCells, icsf_foam_field** out);
/* nComp (1 or 3) and the internal values [nCells][nComp] (either may be NULL) */
int icsf_foam_field_internal(const icsf_foam_field* f, int* nComp, double* values);
/* boundaryField/<patch>/type */
int icsf_foam_field_patch_type(const icsf_foam_field* f, const char* patch, char* type, int len);
/* boundaryField/<patch>/<key> (value, p0, T0, tangentialVelocity,
 * inletValue, ...) expanded to nFaces values [nFaces][nComp]; uniform,
 * nonuniform and $internalField; ICSF_ERR_ARG when absent. */
int icsf_foam_field_patch_entry(const icsf_foam_field* f, const char* patch, const char* key, int nFaces,
                                double* values);
int icsf_foam_field_free(icsf_foam_field* f);
/* Write a volScalarField (nComp 1) or volVectorField (nComp 3): internal
 * values nonuniform, per patch its type, the entry text patchEntries[p]
 * (complete "keyword value;" lines, e.g. p0 and gamma of totalPressure;
 * array or element may be NULL) and, where patchValues[p] is not NULL, a
 * nonuniform value entry. */
int icsf_foam_field_write(const char* path, const char* name, const char* dimensions, int nComp, int nCells,
                          const double* internal, int nPatches, const char* const* patchNames,
                          const char* const* patchTypes, const int* patchSizes, const double* const* patchValues,
                          const char* const* patchEntries);

/* ---- harmonic balance (dbnsFullyImplicitHBFoam) ----------------------
 * Spectral operator D [nInstants][nInstants] (row-major) of
 * HBZone::updateHBOperators (HBZone.C:270-356) for the omega list of
 * HBZone::setOmegaList (HBZone.C:46-91): 0, +k w_i (k = 1..harmonics_i),
 * then the negative frequencies in reverse order; frequencies are angular
 * (rad/s).  Instants (HBZoneList::setInstants, HBZoneList.C:253-346):
 * t_n = n selectedPeriod / nInstants, or, for selectedPeriod <= 0, the
 * period in [T0, 5 T0] with the smallest condition number.  instants
 * [nInstants] and condition may be NULL. */
int icsf_hb_operator(int nFreq, const double* frequencies, const int* harmonics, int nInstants,
                     double selectedPeriod, double* instants, double* D, double* condition);

/* HB case: nInstants instances of the case (same dictionaries) on one
 * mesh, coupled by D (the coupledMatrix of 2 nInstants scalars and
 * nInstants vectors built in dbnsFullyImplicitHBFoam/outerLoop.H:91-205).
 * fvSolution selects GMRES (double precision) or FGMRES (precision double,
 * or mixed: FP64 outer FGMRES around an FP32 inner GMRES(innerDirections),
 * or around the FP32 LU-SGS alone with innerDirections 0), both with the
 * LUSGS preconditioner. */
int icsf_hb_create(icsf_mesh* mesh, const char* thermo, const char* schemes, const char* solution, int nInstants,
                   const double* D, icsf_hb** out);
/* Post-processing of reconstructHBFoam (applications/solvers/
 * reconstructHBFoam/reconstructHBFoam.C:79-85, HBZone::reconstruct,
 * HBZoneTemplates.C:328-386): the weights w[m] = Re(sum_n exp(i omega_n t)
 * E[n][m]), E = pinv(E^-1) of the instants (HBZone.C:292), of the Fourier
 * reconstruction at time t (frequencies and harmonics as
 * icsf_hb_operator); then the fields at t, sum_m w[m] field_m for rho, U,
 * T, p of every cell ([nCells], U [nCells][3]; any may be NULL). */
int icsf_hb_reconstruct_weights(int nFreq, const double* frequencies, const int* harmonics, int nInstants,
                                const double* instants, double time, double* weights);
int icsf_hb_reconstruct(icsf_hb* h, const double* weights, double* rho, double* U, double* T, double* p);
int icsf_hb_destroy(icsf_hb* h);
/* Per-instant boundary conditions, initial boundary values and fields (as
 * icsf_set_boundary / icsf_set_boundary_values / icsf_set_fields /
 * icsf_get_fields). */
int icsf_hb_set_boundary(icsf_hb* h, int instant, int patch, int pType, int UType, int TType, const double* pValue,
                         const double* UValue, const double* TValue);
int icsf_hb_set_boundary_values(icsf_hb* h, int instant, int patch, const double* pValue, const double* UValue,
                                const double* TValue);
int icsf_hb_set_fields(icsf_hb* h, int instant, const double* rho, const double* U, const double* T);
int icsf_hb_get_fields(icsf_hb* h, int instant, double* rho, double* U, double* T, double* p, double* W);
/* b [nInstants][5][nCells]: R V of each instance plus the HB source
 * -V sum_K D_JK W_K (residualsUpdate.H:72-74, HBZoneTemplates.C:39-93). */
int icsf_hb_residual(icsf_hb* h, double* b);
/* Instance Jacobians plus V D_JK I blocks (outerLoop.H:155-205,
 * HBZone.C:435-518); LU-SGS diagonal over all instances (lusgs.C:50-125). */
int icsf_hb_assemble_jacobian(icsf_hb* h);
/* y = A x over [nInstants][5][nCells] vectors (coupledMatrix::matrixMul,
 * coupledMatrix.C:66-123). */
int icsf_hb_matrix_mul(icsf_hb* h, const double* x, double* y);
/* r <- LU-SGS(r) (lusgs::precondition, lusgs.C:220-382). */
int icsf_hb_precondition(icsf_hb* h, double* r);
/* GMRES increment for b at the current states (coupledMatrix::solveForIncr,
 * outerLoop.H:213-219); residuals [nInstants][10] may be NULL. */
int icsf_hb_solve_for_incr(icsf_hb* h, const double* b, double* dW, double* residuals, int* nIter);
/* nSteps HB pseudo-steps (dbnsFullyImplicitHBFoam/outerLoop.H:1-230 +
 * updateFields.H per instance); residuals [nSteps][nInstants][10] or NULL.
 * As icsf_pseudo_steps, the residuals and the matrix refer to the states
 * before the last update: the matrix calls need icsf_hb_residual and
 * icsf_hb_assemble_jacobian again. */
int icsf_hb_pseudo_steps(icsf_hb* h, int nSteps, double* residuals);
int icsf_hb_synchronize(icsf_hb* h);

#ifdef __cplusplus
}
#endif

#endif /* ICSF_H */
