// hex operators, 5 modes per direction.
#define OPS_NM 5
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
