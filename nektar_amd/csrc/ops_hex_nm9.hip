// hex operators, 9 modes per direction.
#define OPS_NM 9
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
