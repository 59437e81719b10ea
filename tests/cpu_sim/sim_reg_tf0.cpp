// Register tiles of transform type 0 on the CPU model (see sim_reg_tf.inc).
#define SIM_TF 0
#define SIM_FN sim_reg_tf0
#define SIM_FN_ROWS sim_rows_tf0
#include "sim_reg_tf.inc"
