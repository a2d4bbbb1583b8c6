/* Heat-conduction oracle: declarations (heat_oracle.c). TEST INFRASTRUCTURE ONLY. */
#ifndef FALCON_HEAT_ORACLE_H
#define FALCON_HEAT_ORACLE_H
#endif
