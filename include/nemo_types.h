#ifndef NEMO_TYPES_H
#define NEMO_TYPES_H

/* Status codes, backend ids and id types: NeMo src/nemo/types.h:31-71,
 * same values. */

#include <stdint.h>

#define NEMO_OK 0
#define NEMO_CUDA_INVOCATION_ERROR 1   /* a HIP runtime call failed */
#define NEMO_CUDA_ASSERTION_FAILURE 2
#define NEMO_CUDA_MEMORY_ERROR 3       /* a device allocation failed */
#define NEMO_CUDA_ERROR 4              /* other device error (e.g. a step barrier timed out) */
#define NEMO_API_UNSUPPORTED 5
#define NEMO_INVALID_INPUT 6
#define NEMO_BUFFER_OVERFLOW 7
#define NEMO_BUFFER_UNDERFLOW 8
#define NEMO_LOGIC_ERROR 9
#define NEMO_ALLOCATION_ERROR 10
#define NEMO_MPI_ERROR 11
#define NEMO_DL_ERROR 12
#define NEMO_IO_ERROR 13
#define NEMO_UNKNOWN_ERROR 14

enum {
	NEMO_BACKEND_CUDA,   /* the GPU backend slot; this library's HIP backend */
	NEMO_BACKEND_CPU
};

typedef unsigned backend_t;
typedef unsigned long long cycle_t;
typedef uint64_t synapse_id;

#endif
