// pvglu_k4.hip -- numeric kernels at instance width 4 (one translation
// unit per width, so the widths compile in parallel).
#include "pvglu_kernels.h"

template struct pvglu::Launch<4>;
