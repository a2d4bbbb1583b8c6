#include "falcon_oracle.h"
