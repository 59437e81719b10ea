// Register tiles of transform type 1 on the CPU model (see sim_reg_tf.inc).
#define SIM_TF 1
#define SIM_FN sim_reg_tf1
#define SIM_FN_ROWS sim_rows_tf1
#include "sim_reg_tf.inc"
