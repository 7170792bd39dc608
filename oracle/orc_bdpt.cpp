#include "orc_render.h"
