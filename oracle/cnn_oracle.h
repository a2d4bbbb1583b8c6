/* CNN oracle: declarations (cnn_oracle.c). TEST INFRASTRUCTURE ONLY. */
#ifndef FALCON_CNN_ORACLE_H
#define FALCON_CNN_ORACLE_H
#endif
