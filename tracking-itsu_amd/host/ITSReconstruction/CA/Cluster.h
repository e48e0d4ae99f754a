// Cluster.h -- forwards to Event.h (the reference keeps one type per header).
#pragma once
#include "Event.h"
