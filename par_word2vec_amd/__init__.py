"""MI355X-native PAR-Word2Vec attraction-repulsion SGNS trainer.

Host side (C++, libpar_host.so): tokenizer, vocabulary, subsampling into a
sentence CSR, the reference's exact random schedule, synthetic corpora,
embedding writer.  Device side (HIP for gfx950, libpar_sgns.so): the
deterministic single-wavefront executor, the plain and persistent Hogwild
kernels (fp32 / bf16 tables), the window loss-forward and adjoint kernels.
Python only binds the two C-ABIs (include/par_host.h, include/par_sgns.h)."""
from . import host
from .device import SgnsDevice
from .trainer import Options, ParWord2Vec

__all__ = ["host", "SgnsDevice", "Options", "ParWord2Vec"]
