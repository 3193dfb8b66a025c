// The learned dispatcher's network shape and its device-side entry used by
// the batch engine.
#pragma once
#include <stddef.h>
#include <stdint.h>

#define FF_MLP_IN 8
#define FF_MLP_HIDDEN 64
#define FF_MLP_MAXK 32

// Candidate count of the loaded network (0: none loaded).
int ff_mlp_ready(void);
// Inference on n device feature rows (n x FF_MLP_IN floats): argmin over the
// candidates allowed by each row's mask bits into dchoice, outputs into dy
// when not null; with margin > 0, candidate dflt stays unless another is
// predicted faster by more than margin (log-time). Queued on the library
// stream.
int ff_mlp_infer_dev(const float *dx, size_t n, const uint32_t *dmask, float *dy, int *dchoice, float margin,
                     int dflt);

// The dispatcher's margin: deviate from the default schedule only for a
// predicted gain above 2%.
#define FF_MLP_MARGIN 0.02f
