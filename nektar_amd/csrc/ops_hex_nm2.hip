// hex operators, 2 modes per direction.
#define OPS_NM 2
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
