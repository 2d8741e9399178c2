// pvglu_k2.hip -- numeric kernels at instance width 2 (one translation
// unit per width, so the widths compile in parallel).
#include "pvglu_kernels.h"

template struct pvglu::Launch<2>;
