// pvglu_k64.hip -- numeric kernels at instance width 64 (one translation
// unit per width, so the widths compile in parallel).
#include "pvglu_kernels.h"

template struct pvglu::Launch<64>;
