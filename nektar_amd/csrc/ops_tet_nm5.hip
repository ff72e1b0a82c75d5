// tet operators, 5 modes per direction.
#define OPS_NM 5
#define OPS_SHAPE_TET
#include "ops_inst.inc"
