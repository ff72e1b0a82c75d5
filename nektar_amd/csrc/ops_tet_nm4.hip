// tet operators, 4 modes per direction.
#define OPS_NM 4
#define OPS_SHAPE_TET
#include "ops_inst.inc"
