// Cell.h -- forwards to Road.h (the reference keeps one type per header).
#pragma once
#include "Road.h"
