// tet operators, 8 modes per direction.
#define OPS_NM 8
#define OPS_SHAPE_TET
#include "ops_inst.inc"
