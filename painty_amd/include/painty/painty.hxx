// Everything of the drop-in interface.
#pragma once

#include "painty/core/KubelkaMunk.hxx"
#include "painty/core/Spline.hxx"
#include "painty/mixer/PaintMixer.hxx"
#include "painty/mixer/Serialization.hxx"
#include "painty/renderer/FootprintBrush.hxx"
#include "painty/renderer/Renderer.hxx"
#include "painty/renderer/TextureBrush.hxx"
