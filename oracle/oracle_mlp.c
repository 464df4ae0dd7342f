/* placeholder: the MLP oracle is added with the surrogate */
#include "sirius_oracle.h"
