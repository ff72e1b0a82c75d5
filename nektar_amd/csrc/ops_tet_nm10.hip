// tet operators, 10 modes per direction.
#define OPS_NM 10
#define OPS_SHAPE_TET
#include "ops_inst.inc"
