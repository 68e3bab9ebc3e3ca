/* learned-tau MLP reference: filled in below */
#include "plis_oracle.h"

