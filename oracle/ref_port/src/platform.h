/* fastfib.c includes "src/platform.h" and the modules include "platform.h";
   both names lead to the one Linux layer (see ../platform.h). */
#pragma once
#include "platform.h"
