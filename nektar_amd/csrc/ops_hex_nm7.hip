// hex operators, 7 modes per direction.
#define OPS_NM 7
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
