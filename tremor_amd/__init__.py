"""MI355X-native exact template matching (z-normalised Euclidean distance /
Pearson correlation, kNN and threshold search) behind a C-ABI.

    from tremor_amd.capi import Context
    with Context(precision="bf16x3") as ctx:
        ctx.load_waveform(x)
        pos, d, stats = ctx.knn(templates, k=10)
"""
