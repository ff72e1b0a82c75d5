// hex operators, 10 modes per direction.
#define OPS_NM 10
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
