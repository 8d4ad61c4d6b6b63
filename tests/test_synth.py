"""CPU: the synthetic workload images are pinned by SHA-256, so every run
(tests, bench, CPU baseline) encodes exactly the same pixels: the 512x512
grayscale case (seed 0) and the 4096x4096 RGB case (channel seeds 1, 2, 3,
luminance walks seed 4)."""
import hashlib

from roger_amd import synth

SHA_GRAY_512 = "8453f9a4e8578697d706c2956f3691338f18e47e8d5e880337e0106a7b6e020f"
SHA_RGB_4096 = "97269412216cc391cce181a4ee52a09ad8f51b6940b15feda36c8ffb1ac7ed90"


def test_gray_512_is_pinned():
    img = synth.image(512, 512, 1, 0)
    assert img.shape == (512, 512)
    assert hashlib.sha256(img.tobytes()).hexdigest() == SHA_GRAY_512


def test_rgb_4096_is_pinned():
    img = synth.rgb(4096, 4096, 1)
    assert img.shape == (4096, 4096, 3)
    assert hashlib.sha256(img.tobytes()).hexdigest() == SHA_RGB_4096
