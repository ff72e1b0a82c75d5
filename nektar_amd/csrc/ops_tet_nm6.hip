// tet operators, 6 modes per direction.
#define OPS_NM 6
#define OPS_SHAPE_TET
#include "ops_inst.inc"
