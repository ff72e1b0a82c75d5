// tet operators, 7 modes per direction.
#define OPS_NM 7
#define OPS_SHAPE_TET
#include "ops_inst.inc"
