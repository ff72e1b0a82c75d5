// hex operators, 8 modes per direction.
#define OPS_NM 8
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
