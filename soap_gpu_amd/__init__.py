"""soap_gpu_amd: MI355X-native (gfx950 HIP) engine for SOAP-style stellar activity spectra."""
