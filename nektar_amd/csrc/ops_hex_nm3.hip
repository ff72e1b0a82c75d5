// hex operators, 3 modes per direction.
#define OPS_NM 3
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
