// tet operators, 2 modes per direction.
#define OPS_NM 2
#define OPS_SHAPE_TET
#include "ops_inst.inc"
