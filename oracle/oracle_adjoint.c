#include "mcint_oracle.h"
