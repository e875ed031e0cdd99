"""HDF5 microstructure input and results output (reader.cpp:227-411, reader.h:88-351)."""


def read_microstructure(path, dataset):
    raise RuntimeError("HDF5 input requires the results/IO module; pass the microstructure array explicitly")
