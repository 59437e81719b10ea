// Register tiles of transform type 2 on the CPU model (see sim_reg_tf.inc).
#define SIM_TF 2
#define SIM_FN sim_reg_tf2
#define SIM_FN_ROWS sim_rows_tf2
#include "sim_reg_tf.inc"
