// Host-side polyMesh: addressing, renumbering, geometry and the device
// addressing tables (cell->face pull lists, sliced-ELL block slots, sweep
// levels). Pure C++; no HIP calls.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace icsf {

enum PatchKind { PATCH_PATCH = 0, PATCH_WALL = 1, PATCH_EMPTY = 2, PATCH_CYCLIC = 3 };

// Slice height of the sliced-ELL block storage: one wavefront of rows.
constexpr int SLICE = 64;

struct HostMesh {
    int nPoints = 0, nCells = 0, nFaces = 0, nInternal = 0;
    std::vector<double> points;              // [nPoints*3]
    std::vector<int> faceOff, faceVerts;     // CSR polygon list
    std::vector<int> owner, neighbour;       // [nFaces], [nInternal]
    std::vector<int> patchStart, patchSize, patchKind;
    std::vector<int> patchNbr;               // [nPatches] paired cyclic patch, -1
    std::vector<double> patchT;              // [nPatches][9] cyclic forwardT (row-major)

    // geometry (OpenFOAM primitiveMesh / surfaceInterpolation definitions)
    std::vector<double> Sf, magSf, Cf;       // [nFaces*3], [nFaces], [nFaces*3]
    std::vector<double> C, V;                // [nCells*3], [nCells]
    std::vector<double> weights;             // [nFaces] (1 on boundary faces)
    std::vector<double> deltaCoeffs;         // [nFaces]
    std::vector<double> nonOrthDeltaCoeffs;  // [nFaces]
    std::vector<double> corrVecs;            // [nFaces*3] (0 on boundary faces)

    // derived addressing
    std::vector<int> cellFaceOff, cellFaces; // faces of each cell sorted by face
                                             // index, empty-patch faces omitted
    std::vector<int> faceIsEmpty;            // [nFaces - nInternal]

    // cyclic faces, numbered k = 0.. in face order
    std::vector<int> cycOf;                  // [nFaces - nInternal] k, or -1
    std::vector<int> cycFace;                // [nCyc] face
    std::vector<int> cycCell;                // [nCyc] cell on the far side (owner of the paired face)
    std::vector<int> cycPair;                // [nCyc] paired face
    std::vector<double> cycR;                // [nCyc][9] forwardT of the face's patch
    std::vector<double> cycDelta;            // [nCyc][3] cyclicFvPatch::delta()
    std::vector<int> cycRowOff, cycRowK;     // cell -> its cyclic faces (face order)
    bool twoD = false;                       // has an empty patch
    int emptyDir = -1;                       // direction removed by the empty patch

    int patchOf(int face) const;
};

// Build a HostMesh from raw arrays (copies), compute geometry + pull lists.
std::string buildHostMesh(HostMesh& m, int nPoints, const double* points, int nFaces,
                          const int* faceOff, const int* faceVerts, int nInternal,
                          const int* owner, const int* neighbour, int nPatches,
                          const int* patchStart, const int* patchSize, const int* patchKind,
                          const int* patchNbr = nullptr, const double* patchT = nullptr);

void computeGeometry(HostMesh& m);
// Pairing and geometry of cyclic faces (after computeGeometry); error text
// or empty.
std::string computeCyclic(HostMesh& m);
void buildCellFaces(HostMesh& m);

// --- renumbering (integer work; must be reproducible bit-for-bit) -------
// cellCells in internal-face order (OpenFOAM primitiveMesh::calcCellCells).
void cellCells(int nCells, int nInternal, const int* owner, const int* neighbour,
               std::vector<int>& off, std::vector<int>& adj);
// The same with the connections across coupled (cyclic) faces appended to
// each cell's list in boundary-face order (decompositionMethod::calcCellCells,
// which OpenFOAM's renumberMesh uses): nCoupled faces with the cell on each
// side.
void cellCellsCoupled(int nCells, int nInternal, const int* owner, const int* neighbour, int nCoupled,
                      const int* cOwner, const int* cNbr, std::vector<int>& off, std::vector<int>& adj);
// Cuthill-McKee band compression (OpenFOAM meshTools::bandCompression),
// reversed: newToOld.
void rcmOrder(int nCells, int nInternal, const int* owner, const int* neighbour,
              std::vector<int>& newToOld, int nCoupled = 0, const int* cOwner = nullptr,
              const int* cNbr = nullptr);
// Greedy colouring in cell-index order.
int greedyColour(int nCells, int nInternal, const int* owner, const int* neighbour,
                 std::vector<int>& colour);
// Stable partition of cells by colour: newToOld.
void colourMajorOrder(const std::vector<int>& colour, int nColours, std::vector<int>& newToOld);
// Apply a cell permutation; internal faces are re-oriented (owner<neighbour,
// flipped faces reversed) and re-sorted upper-triangular, boundary faces keep
// their order.  faceNewToOld[newFace] = oldFace (negative-1 encoded if flipped:
// -(old+1)).
void renumberMesh(int nCells, int nFaces, int nInternal, const int* faceOff,
                  const int* faceVerts, const int* owner, const int* neighbour,
                  const int* newToOld, std::vector<int>& faceVertsOut,
                  std::vector<int>& ownerOut, std::vector<int>& neighbourOut,
                  std::vector<int>& faceNewToOld);

// Sweep levels of the lower-triangular (forward) and upper-triangular
// (backward) dependency graphs in the current numbering.
void sweepLevels(int nCells, int nInternal, const int* owner, const int* neighbour,
                 std::vector<int>& fwdLevel, std::vector<int>& bwdLevel,
                 int& nFwd, int& nBwd);

// blockMesh-style topology of a structured hexahedral block.
void hexBlockTopology(int nx, int ny, int nz, const int* sidePatch, int nPatches, int* faceVerts, int* owner,
                      int* neighbour, int* patchSize);

// Sliced-ELL off-diagonal block layout.
struct SellLayout {
    int nCells = 0, nSlices = 0;
    std::vector<int> sliceWidth;     // K_s
    std::vector<int> slicePtr;       // prefix sums of K_s, size nSlices+1
    std::vector<int> col;            // [slicePtr[nSlices]*SLICE], -1 = padding
    std::vector<int> rowLen;         // [nCells]
    std::vector<int> slotUpper;      // [nInternal]: slot of (owner, neighbour) in owner row
    std::vector<int> slotLower;      // [nInternal]: slot of (neighbour, owner) in neighbour row
    std::vector<int> nLower;         // [nCells]: number of slots with column < row
};
void buildSell(int nCells, int nInternal, const int* owner, const int* neighbour, SellLayout& L);

}  // namespace icsf
