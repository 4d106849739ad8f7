"""spider_amd: MI355X-native Fourier-gridding backprojection (SPIDER `BP 3F` / `BP 32F`)."""
