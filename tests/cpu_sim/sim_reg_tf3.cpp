// Register tiles of transform type 3 on the CPU model (see sim_reg_tf.inc).
#define SIM_TF 3
#define SIM_FN sim_reg_tf3
#define SIM_FN_ROWS sim_rows_tf3
#include "sim_reg_tf.inc"
