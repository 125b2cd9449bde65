#include "esth_internal.h"
