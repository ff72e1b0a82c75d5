// hex operators, 6 modes per direction.
#define OPS_NM 6
#define OPS_SHAPE_HEX
#include "ops_inst.inc"
